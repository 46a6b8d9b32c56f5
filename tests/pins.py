"""Closed-form cases that pin the oracle (and, at full size, the GPU core)
without the reference: an affine grid carrying cubic wavefields, on which
every 5-point MacCormack operator is exact and the semi-discrete system is
nilpotent, so one RK4 step equals the exact solution.

Why it pins: the fields are cubics P_c(u) of u = n.r / L with a generic unit
vector n and a non-orthogonal affine grid r = A (i, j, k), so every metric
term (xi_x .. zeta_z, all nonzero), every derivative direction and every
term of the momentum and Hooke sums contributes; a dropped term, a sign, a
transposed metric index or a wrong RK coefficient changes the step by O(1)
of the increment.
"""
from __future__ import annotations

import numpy as np

# component order Vx,Vy,Vz,Txx,Tyy,Tzz,Tyz,Txz,Txy
TIDX = {(0, 0): 3, (1, 1): 4, (2, 2): 5, (1, 2): 6, (2, 1): 6, (0, 2): 7, (2, 0): 7, (0, 1): 8, (1, 0): 8}

VP, VS, RHO = 6000.0, 3464.0, 2670.0


def affine_matrix(dh: float, seed: int = 11) -> np.ndarray:
    """dx/d(i,j,k): a sheared, non-orthogonal grid with spacing ~dh and det > 0."""
    rng = np.random.Generator(np.random.PCG64(seed))
    A = dh * (np.eye(3) + 0.25 * rng.uniform(-1.0, 1.0, size=(3, 3)))
    if np.linalg.det(A) < 0:
        A[:, 0] *= -1.0
    return A


def affine_scenario(ni: int, nj: int, nk: int, A: np.ndarray, dt: float, pml: int = 0, free_top: bool = False):
    """A fault-free scenario on the grid r = A (i, j, k) with the homogeneous
    medium of md_t.cu:396-398."""
    from cgfd import configs
    from cgfd.model import Scenario
    ii = np.arange(ni, dtype=np.float64)[None, None, :]
    jj = np.arange(nj, dtype=np.float64)[None, :, None]
    kk = np.arange(nk, dtype=np.float64)[:, None, None]
    r = [A[a, 0] * ii + A[a, 1] * jj + A[a, 2] * kk for a in range(3)]
    x, y, z = (np.ascontiguousarray(np.broadcast_to(c, (nk, nj, ni)), np.float32) for c in r)
    lam, mu, rho = configs.homogeneous_media((nk, nj, ni), VP, VS, RHO)
    layers = [[pml, pml], [pml, pml], [pml, 0 if free_top else pml]]
    return Scenario(ni=ni, nj=nj, nk=nk, x=x, y=y, z=z, lam=lam, mu=mu, rho=rho, free_top=free_top,
                    pml_layers=layers, faults=[], dt=dt, nt_total=4, name="affine")


def cubic_coefs(seed: int = 5) -> np.ndarray:
    """(9, 4) coefficients of P_c(u) = sum_m C[c, m] u^m: velocities ~1 m/s,
    stresses ~rho*Vp*1 m/s."""
    rng = np.random.Generator(np.random.PCG64(seed))
    C = rng.uniform(-1.0, 1.0, size=(9, 4))
    C[3:] *= RHO * VP
    return C


def unit_vector(seed: int = 3) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    n = rng.normal(size=3)
    return n / np.linalg.norm(n)


def _dpoly(c: np.ndarray) -> np.ndarray:
    out = np.zeros_like(c)
    out[..., :-1] = c[..., 1:] * np.arange(1, c.shape[-1])
    return out


def rhs_coefs(C: np.ndarray, n: np.ndarray, L: float, lam: float, mu: float, rho: float) -> np.ndarray:
    """Time derivative of the coefficient array under the isotropic elastic
    velocity-stress equations, fields depending on u = n.r / L only:
    dV_i/dt = (1/rho) sum_j d_j T_ij,  dT_ij/dt = lam delta_ij div V + mu (d_i V_j + d_j V_i)."""
    dC = _dpoly(C) / L
    out = np.zeros_like(C)
    for i in range(3):
        out[i] = sum(n[j] * dC[TIDX[(i, j)]] for j in range(3)) / rho
    div = sum(n[k] * dC[k] for k in range(3))
    for (i, j), c in ((( 0, 0), 3), ((1, 1), 4), ((2, 2), 5), ((1, 2), 6), ((0, 2), 7), ((0, 1), 8)):
        out[c] = mu * (n[i] * dC[j] + n[j] * dC[i]) + (lam * div if i == j else 0.0)
    return out


def exact_step(C: np.ndarray, n: np.ndarray, L: float, dt: float, nsteps: int = 1) -> np.ndarray:
    """exp(dt Lop)^nsteps applied to the coefficients (Lop^4 = 0 on cubics, so
    the truncated series is exact and equals classical RK4)."""
    lam = VP * VP * RHO - 2.0 * VS * VS * RHO
    mu = VS * VS * RHO
    out = C.copy()
    for _ in range(nsteps):
        term = out.copy()
        acc = out.copy()
        for m in range(1, 5):
            term = rhs_coefs(term, n, L, lam, mu, RHO) * (dt / m)
            acc = acc + term
        out = acc
    return out


def eval_fields(C: np.ndarray, u: np.ndarray, dtype=np.float64) -> np.ndarray:
    """(9, *u.shape) field values."""
    out = np.empty((9,) + u.shape, dtype)
    for c in range(9):
        v = np.zeros(u.shape, np.float64)
        for m in range(C.shape[1] - 1, -1, -1):
            v = v * u + C[c, m]
        out[c] = v
    return out


def u_of_grid(ni: int, nj: int, nk: int, A: np.ndarray, n: np.ndarray, ext: bool = True):
    """u = n.r / L on the extended (ghost-padded) grid, centred on the domain
    middle; L is half the domain extent along n (so |u| <= ~1)."""
    g = 3 if ext else 0
    ii = np.arange(-g, ni + g, dtype=np.float64)[None, None, :]
    jj = np.arange(-g, nj + g, dtype=np.float64)[None, :, None]
    kk = np.arange(-g, nk + g, dtype=np.float64)[:, None, None]
    c = np.array([(ni - 1) / 2.0, (nj - 1) / 2.0, (nk - 1) / 2.0])
    w = n @ A  # d(n.r)/d(i,j,k)
    L = 0.5 * (abs(w[0]) * ni + abs(w[1]) * nj + abs(w[2]) * nk)
    u = (w[0] * (ii - c[0]) + w[1] * (jj - c[1]) + w[2] * (kk - c[2])) / L
    return u, L
