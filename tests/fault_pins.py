"""Closed-form cases that pin the split-node fault path of the oracle (and,
through the same cases, the GPU core).

* Uniform stress drop (``configs.stress_drop``): on a fault covering the
  plane with mu_s = mu_d and a shear traction dtau above the dynamic
  strength, every unwelded node slips from the first step and, until waves
  from the welded border arrive, the solution is the 1-D one: slip rate
  2 dtau / (rho Vs), particle velocity +-dtau / (rho Vs) on the two sides
  (zero on the plane, the side average), shear stress change -dtau behind
  the S front, shear traction mu_d sigma_n, rupture time dt.
* Uniform strain rate: a velocity field V = G (r - r_c) linear in space, no
  stress, and a locked fault. The exact solution keeps V and grows a uniform
  stress sigma = t C:sym(G); every operator of the scheme, including the
  split-node velocity and stress rows (both methods, their matrices, the T1
  rate of method 2) and the free-surface rows, is exact on it, so one step
  reproduces it to rounding wherever the zero ghosts have not reached.
"""
from __future__ import annotations

import numpy as np

from pins import RHO, TIDX, VP, VS

MPA = 1.0e6


def stress_drop_expect(sc):
    """Closed-form slip rate and side velocity of a stress_drop scenario."""
    mu = float(sc.mu.flat[0])
    rho = float(sc.rho.flat[0])
    vs = np.sqrt(mu / rho)
    f = sc.faults[0].fields
    sn = -float(f["Tx"].flat[0])
    mud = float(f["mu_d"].flat[0])
    shear = float((f["Ty"] if sc.free_top else f["Tz"]).flat[0])
    dtau = shear - mud * sn
    return {"rate": 2.0 * dtau / (rho * vs), "side": dtau / (rho * vs), "dtau": dtau, "strength": mud * sn,
            "sn": sn}


def stress_drop_window(sc):
    """(j, k) physical node ranges of the fault plane and the last step at
    which waves from the welded border have not reached them (P waves at
    Vp, with a margin of two steps)."""
    dh = float(sc.y[0, 1, 0] - sc.y[0, 0, 0])
    m = sc.united_margin
    jc = sc.nj // 2
    js = np.arange(jc - 5, jc + 6)
    if sc.free_top:
        ks = np.arange(sc.nk - 10, sc.nk)
        dist = min(js[0] - m, sc.nj - m - 1 - js[-1], ks[0] - m)
    else:
        kc = sc.nk // 2
        ks = np.arange(kc - 5, kc + 6)
        dist = min(js[0] - m, sc.nj - m - 1 - js[-1], ks[0] - m, sc.nk - m - 1 - ks[-1])
    vp = float(np.sqrt((sc.lam.flat[0] + 2 * sc.mu.flat[0]) / sc.rho.flat[0]))
    last = int(dist * dh / (vp * sc.dt)) - 2
    return js, ks, last


def check_stress_drop(fault_out, fault_state, wave, sc, step, tol_rate=5e-3, tol_slip=1.5e-2, tol_wave=1e-2):
    """Compares one run of a stress_drop scenario after `step` steps with the
    closed form; returns a dict of relative errors (asserting is the
    caller's). fault_out [11][nz][ny], fault_state [9][2][nz][ny], wave
    [9][nz][ny][nx], all ghost-inclusive."""
    e = stress_drop_expect(sc)
    js, ks, _ = stress_drop_window(sc)
    J, K = np.meshgrid(js + 3, ks + 3)
    comp = 1 if sc.free_top else 2                      # slip direction: y (strike) or z (dip)
    s_cmp = 4 if sc.free_top else 5                     # Vs1 or Vs2 output
    t_cmp = 1 if sc.free_top else 2                     # Ts1 or Ts2 output
    t = step * sc.dt
    out = {}
    out["rate"] = float(np.abs(fault_out[3][K, J] / e["rate"] - 1.0).max())
    out["rate_dir"] = float(np.abs(np.abs(fault_out[s_cmp][K, J]) / e["rate"] - 1.0).max())
    out["slip"] = float(np.abs(fault_out[6][K, J] / (e["rate"] * t) - 1.0).max())
    out["init_t0"] = float(np.abs(fault_out[10][K, J] / sc.dt - 1.0).max())
    out["strength"] = float(np.abs(np.abs(fault_out[t_cmp][K, J]) / e["strength"] - 1.0).max())
    out["normal"] = float(np.abs(fault_out[0][K, J] / -e["sn"] - 1.0).max())
    vm, vp_ = fault_state[comp, 0][K, J], fault_state[comp, 1][K, J]
    out["side"] = float(max(np.abs(np.abs(vm) / e["side"] - 1.0).max(), np.abs(np.abs(vp_) / e["side"] - 1.0).max()))
    out["side_sign"] = float(np.sign(vp_ - vm).min())                  # +1: the plus side moves along +shear
    i0 = sc.faults[0].x_index + 3
    # the wavefield: velocity on the plane (side average) is zero, and within
    # 8 nodes of the plane it is +-dtau/(rho Vs) behind the S front
    out["plane_velocity"] = float(np.abs(wave[comp][K, J, i0]).max() / e["side"])
    vs_nodes = VS * t / float(sc.x[0, 0, 1] - sc.x[0, 0, 0])
    d = np.arange(2, min(9, int(vs_nodes) - 6))
    errs = []
    for dd in d:
        errs.append(np.abs(np.abs(wave[comp][K, J, i0 + dd]) / e["side"] - 1.0).max())
        errs.append(np.abs(np.abs(wave[comp][K, J, i0 - dd]) / e["side"] - 1.0).max())
    out["wave_side"] = float(max(errs))
    # shear stress change behind the front on the plus side: -dtau
    s_idx = 8 if sc.free_top else 7                     # Txy or Txz
    out["wave_stress"] = float(np.abs(wave[s_idx][K, J, i0 + 4] / -e["dtau"] - 1.0).max())
    out["ok"] = (out["rate"] <= tol_rate and out["rate_dir"] <= tol_rate and out["slip"] <= tol_slip
                 and out["init_t0"] <= 1e-6 and out["strength"] <= 1e-4 and out["normal"] <= 1e-5
                 and out["side"] <= tol_rate and out["side_sign"] > 0 and out["plane_velocity"] <= 1e-3
                 and out["wave_side"] <= tol_wave and out["wave_stress"] <= tol_wave)
    return out


# ---------------------------------------------------------------- uniform strain rate
def lame():
    mu = float(np.float32(VS * VS * RHO))
    lam = float(np.float32(VP * VP * RHO - 2.0 * mu))
    return lam, mu


def strain_rate_gradient(A: np.ndarray, free_top: bool, seed: int = 12, scale: float = 1e-3) -> np.ndarray:
    """A velocity gradient G (dV_i/dx_j) whose stress rate C:sym(G) leaves
    the top coordinate plane traction-free when free_top (the grid's top is
    the plane k = const, normal grad zeta)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lam, mu = lame()
    E = rng.uniform(-1.0, 1.0, (3, 3))
    E = 0.5 * (E + E.T)
    if free_top:
        n = np.linalg.inv(A)[2]
        n = n / np.linalg.norm(n)
        t1 = np.cross(n, [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 1.0, 0.0])
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(n, t1)
        R = np.stack([t1, t2, n], axis=1)           # local frame (t1, t2, n) as columns
        El = R.T @ E @ R
        El[0, 2] = El[2, 0] = El[1, 2] = El[2, 1] = 0.0
        El[2, 2] = -lam / (lam + 2.0 * mu) * (El[0, 0] + El[1, 1])
        E = R @ El @ R.T
    W = rng.uniform(-1.0, 1.0, (3, 3))
    W = 0.5 * (W - W.T)
    return scale * (E + W)


def stress_rate(G: np.ndarray) -> np.ndarray:
    lam, mu = lame()
    E = 0.5 * (G + G.T)
    return lam * np.trace(E) * np.eye(3) + 2.0 * mu * E


def strain_rate_scenario(imethod: int, A: np.ndarray, ni: int = 40, nj: int = 40, nk: int = 32, dt: float = 0.004,
                         free_top: bool = True):
    """A locked fault on the plane i = ni/2 of the affine grid r = A (i,j,k),
    homogeneous medium, no PML (zero ghosts), welded margin 4."""
    import pins
    from cgfd import configs
    from cgfd.model import FaultSpec
    sc = pins.affine_scenario(ni, nj, nk, A, dt, pml=0, free_top=free_top)
    f = configs._fault_fields(nj, nk)
    f["mu_s"][:] = 1000.0
    f["mu_d"][:] = 1000.0
    f["Dc"][:] = 0.4
    f["Tx"][:] = -100.0 * MPA
    sc.faults = [FaultSpec(x_index=ni // 2, grid=(0, nj - 1, 0, nk - 1), fields=f)]
    sc.imethod = imethod
    sc.united_margin = 4
    sc.nt_total = 10
    sc.name = "strain_rate"
    return sc


def strain_rate_state(sc, A: np.ndarray, G: np.ndarray):
    """Initial wavefield [9][nz][ny][nx] (zero ghosts), fault state
    [9][2][nz][ny], method-2 T1 recurrence [6][nz][ny], and the exact state
    after one step (wavefield, fault state)."""
    nz, ny, nx = sc.shape_ext
    i0 = sc.faults[0].x_index + 3
    ii = np.arange(nx, dtype=np.float64)[None, None, :] - 3
    jj = np.arange(ny, dtype=np.float64)[None, :, None] - 3
    kk = np.arange(nz, dtype=np.float64)[:, None, None] - 3
    c = np.array([(sc.ni - 1) / 2.0, (sc.nj - 1) / 2.0, (sc.nk - 1) / 2.0])
    r = [A[a, 0] * (ii - c[0]) + A[a, 1] * (jj - c[1]) + A[a, 2] * (kk - c[2]) for a in range(3)]
    V = [sum(G[a, b] * r[b] for b in range(3)) for a in range(3)]
    w0 = np.zeros((9, nz, ny, nx))
    for a in range(3):
        w0[a] = np.broadcast_to(V[a], (nz, ny, nx))
    phys = np.zeros((nz, ny, nx), bool)
    phys[3:-3, 3:-3, 3:-3] = True
    w0[:, ~phys] = 0.0
    Sd = stress_rate(G)
    w1 = w0.copy()
    for (a, b), cidx in TIDX.items():
        w1[cidx][phys] = sc.dt * Sd[a, b]
    Ai = np.linalg.inv(A)                        # rows: grad xi, grad eta, grad zeta
    jac = np.linalg.det(A)
    f0 = np.zeros((9, 2, nz, ny))
    vf = [np.broadcast_to(V[a], (nz, ny, nx))[:, :, i0] for a in range(3)]
    for a in range(3):
        f0[a, 0] = vf[a]
        f0[a, 1] = vf[a]
    f1 = f0.copy()
    T2 = jac * Sd @ Ai[1]
    T3 = jac * Sd @ Ai[2]
    for a in range(3):
        f1[3 + a] = sc.dt * T2[a]
        f1[6 + a] = sc.dt * T3[a]
    t1 = np.zeros((6, nz, ny))
    hT1 = jac * Sd @ Ai[0]
    for a in range(3):
        t1[a] = hT1[a]
    return w0, f0, t1, w1, f1


def strain_rate_errors(w, f, w1, f1, sc, clean: int = 14, top_skip: int = 0):
    """Relative max errors of one step against the exact state over nodes
    at least `clean` nodes from the zero-ghost boundaries (the top is
    traction-free and exact when free_top), leaving out the top `top_skip`
    rows. Returns (wave velocity, wave stress, fault-state velocity,
    fault-state traction)."""
    g = 3
    kt = (sc.nk if sc.free_top else sc.nk - clean) - top_skip
    sl = (slice(g + clean, g + kt), slice(g + clean, g + sc.nj - clean), slice(g + clean, g + sc.ni - clean))
    fs = (slice(g + clean, g + kt), slice(g + clean, g + sc.nj - clean))
    ev = np.abs(w[:3][(slice(None),) + sl] - w1[:3][(slice(None),) + sl]).max() / np.abs(w1[:3][(slice(None),) + sl]).max()
    es = np.abs(w[3:][(slice(None),) + sl] - w1[3:][(slice(None),) + sl]).max() / np.abs(w1[3:][(slice(None),) + sl]).max()
    fv = np.abs(f[:3][(slice(None), slice(None)) + fs] - f1[:3][(slice(None), slice(None)) + fs]).max() / \
        np.abs(f1[:3][(slice(None), slice(None)) + fs]).max()
    ft = np.abs(f[3:][(slice(None), slice(None)) + fs] - f1[3:][(slice(None), slice(None)) + fs]).max() / \
        np.abs(f1[3:][(slice(None), slice(None)) + fs]).max()
    return float(ev), float(es), float(fv), float(ft)
