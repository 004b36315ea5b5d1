"""What the trained network adds over the method without it.

The same cut-cell discretisation is (a) solved for grid values by BiCGSTAB
(GridJudge) and (b) used to train the two-phase network (the NBM).  On a finer
evaluation grid the network is evaluated directly while the grid solution has
to be interpolated (trilinear); the network represents the jump across the
interface by construction, the interpolant smears it.  Measured on the 24^3
sphere problem (2000 epochs, 0.2 s of training on an MI355X): network L-inf
0.25 / L2 0.035 against 0.49 / 0.050 for the interpolated grid solve.
"""
import numpy as np
import pytest

from jax_dips import problems


@pytest.mark.gpu
def test_network_beats_interpolated_grid_solve(tmp_path):
    from scipy.interpolate import RegularGridInterpolator

    from jax_dips.domain import mesh
    from jax_dips.judge import GridJudge
    from jax_dips.solvers.poisson import trainer

    prob = problems.sphere()
    n = 24
    x, y, z = problems.grid_axes(n)
    g = mesh.construct(3)[0](x, y, z)
    xe = problems.grid_axes(2 * n - 1)[0]
    ge = mesh.construct(3)[0](xe, xe, xe)
    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"])
    st, solve = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=ge, num_epochs=2000, batch_size=n ** 3,
                        checkpoint_dir=str(tmp_path), results_dir=str(tmp_path),
                        optimizer_dict={"optimizer_name": "custom", "learning_rate": 1e-2,
                                        "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}})
    st2, _, loss = solve(st)
    # training ran: the loss drops by well over two orders of magnitude (where
    # the 2000-step trajectory levels off, ~7e-4 to 1e-3 of the first epoch,
    # moves with the float32 rounding of the reductions)
    print("final / first epoch loss", loss[-1] / loss[0])
    assert loss[-1] < 5e-3 * loss[0]
    Re = np.asarray(ge.R, np.float64)
    exact = prob["exact_fn"](Re.T)
    e_nn = np.abs(st2.solution - exact)
    ug, it, rr = GridJudge(g, g, problems.oracle_callbacks(prob)).solve(rtol=1e-7)
    assert rr <= 1e-6
    e_grid = np.abs(RegularGridInterpolator((x, y, z), ug.reshape(n, n, n))(Re) - exact)
    print("Linf nn", e_nn.max(), "grid", e_grid.max(), "L2 nn", np.sqrt((e_nn ** 2).mean()),
          "grid", np.sqrt((e_grid ** 2).mean()))
    assert e_nn.max() < 0.75 * e_grid.max()
    assert np.sqrt((e_nn ** 2).mean()) < 0.9 * np.sqrt((e_grid ** 2).mean())
