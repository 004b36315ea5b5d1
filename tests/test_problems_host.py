"""The variable-coefficient problems restating the reference's test
configurations (problems.ref_sphere, problems.ref_star; tests/confs/
experiment_configs.py:21-186, :196-412) are consistent with their
manufactured solutions, checked against central finite differences of the
callbacks themselves (h = 1e-4, float64):
* grad phi, grad u-, grad u+, grad mu-, grad mu+ against differences of phi, u, mu;
* f = -div(mu grad u) on each side against differences of the flux mu grad u;
* alpha = u+ - u- and beta = -(mu+ grad u+ - mu- grad u-) . grad phi;
* mu- > mu+ in the box (both configurations exercise the mu- > mu+ ghost values).
"""
import numpy as np
import pytest

from jax_dips import problems

H = 1e-4


def fd_grad(fn, r):
    out = []
    for a in range(3):
        e = np.zeros((3, 1))
        e[a] = H
        out.append((fn(r + e) - fn(r - e)) / (2 * H))
    return np.stack(out)


@pytest.mark.parametrize("name", ["ref_sphere", "ref_star"])
def test_reference_problems_are_consistent(name):
    prob = getattr(problems, name)()
    rng = np.random.default_rng(5)
    r = rng.uniform(-0.95, 0.95, (3, 400))
    r = r[:, np.linalg.norm(r, axis=0) > 0.1]
    for g_fn, fn in (("grad_phi_fn", "phi_raw_fn"), ("grad_exact_m_fn", "exact_m_fn"),
                     ("grad_exact_p_fn", "exact_p_fn"), ("grad_mu_m_fn", "mu_m_fn"), ("grad_mu_p_fn", "mu_p_fn")):
        an = np.stack(np.broadcast_arrays(*prob[g_fn](r)))
        fd = fd_grad(prob[fn], r)
        assert np.max(np.abs(an - fd)) <= 1e-5 * max(1.0, np.max(np.abs(fd))), g_fn
    for side in ("m", "p"):
        mu, gu = prob[f"mu_{side}_fn"], prob[f"grad_exact_{side}_fn"]
        div = 0.0
        for a in range(3):
            e = np.zeros((3, 1))
            e[a] = H
            div = div + (mu(r + e) * gu(r + e)[a] - mu(r - e) * gu(r - e)[a]) / (2 * H)
        f = prob[f"f_{side}_fn"](r)
        assert np.max(np.abs(f + div)) <= 1e-5 * max(1.0, np.max(np.abs(f))), side
    assert np.allclose(prob["alpha_fn"](r), prob["exact_p_fn"](r) - prob["exact_m_fn"](r))
    gp = np.stack(prob["grad_phi_fn"](r))
    flux = prob["mu_p_fn"](r) * np.stack(prob["grad_exact_p_fn"](r)) - prob["mu_m_fn"](r) * np.stack(
        np.broadcast_arrays(*prob["grad_exact_m_fn"](r)))
    assert np.allclose(prob["beta_fn"](r), -(flux * gp).sum(axis=0))
    box = rng.uniform(-1, 1, (3, 5000))
    assert np.all(prob["mu_m_fn"](box) > prob["mu_p_fn"](box))
    assert np.all(np.isfinite(prob["lvl_set_fn"](np.zeros((3, 1)))))


def brute_atoms(c, rad, q, x, sigma=0.5):
    """phi = min over every atom of |x - c_i| - r_i and the full Gaussian charge
    sum: the reference's brute-force scan (examples/biomolecules/geometry.py:10-39)."""
    d2 = ((x.T[:, None, :] - c[None]) ** 2).sum(axis=2)
    phi = (np.sqrt(d2) - rad[None]).min(axis=1)
    chg = (q[None] * np.exp(-d2 / (2 * sigma ** 2))).sum(axis=1) / (2 * np.pi * sigma ** 2) ** 1.5
    return phi, chg


def test_atom_field_cell_list_matches_brute_force():
    """The 27-cell query equals the all-atom minimum wherever it is below cap
    (and is cap elsewhere), and the charge equals the all-atom sum to within the
    Gaussian tails cut beyond one cell; the compressed rows hold the same atoms
    per cell as the padded table."""
    prob = problems.pb_box(n_atoms=300, half_width=8.0, seed=4)
    f = prob["field"]
    rng = np.random.default_rng(1)
    x = rng.uniform(-8.5, 8.5, (3, 4000))
    phi, chg = f.phi(x), f.charge(x)
    bphi, bchg = brute_atoms(f.c, f.r, f.q, x)
    assert np.array_equal(phi, np.minimum(bphi, f.cap))
    assert np.max(np.abs(chg - bchg)) <= 1e-12 * np.max(np.abs(bchg))
    pad = len(f.c)
    for cell in range(f.nc ** 3):
        rows = f.cell_atoms[f.cell_start[cell]:f.cell_start[cell + 1]]
        assert sorted(rows.tolist()) == sorted(a for a in f.table[cell].tolist() if a != pad)
    assert f.cell_start[-1] == pad and np.array_equal(np.sort(f.cell_atoms), np.arange(pad))
    assert np.array_equal(f.atoms5[:, :3], f.c) and np.array_equal(f.atoms5[:, 3], f.r)
