"""Simulation-state containers (jax_dips/solvers/simulation_states.py:12-82)."""
from typing import Callable

from .._jaxmd_modules import dataclasses
from .._jaxmd_modules.util import Array


@dataclasses.dataclass
class PoissonSimStateFn:
    u_0_fn: Callable
    dir_bc_fn: Callable
    phi_fn: Callable
    mu_m_fn: Callable
    mu_p_fn: Callable
    k_m_fn: Callable
    k_p_fn: Callable
    f_m_fn: Callable
    f_p_fn: Callable
    alpha_fn: Callable
    beta_fn: Callable
    nonlinear_op_m: Callable
    nonlinear_op_p: Callable


@dataclasses.dataclass
class PoissonSimState:
    phi: Array
    solution: Array
    dirichlet_bc: Array
    mu_m: Array
    mu_p: Array
    k_m: Array
    k_p: Array
    f_m: Array
    f_p: Array
    alpha: Array
    beta: Array
    grad_solution: Array
    grad_normal_solution: Array
