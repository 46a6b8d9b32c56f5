"""Friction-parameter gradients on the differentiation grid: 160 x 240 x 100
nodes at 100 m, homogeneous medium with 3% seeded noise, free surface,
PML 12, a 12 x 6 km fault from the surface, 32 surface stations plus 8
near the fault, 64 fault receivers around the nucleation patch. The target
run perturbs Dc and the initial shear traction by seeded smooth fields
(about 10%). Over 40 steps, the adjoint gradients of
  J_vs = 1/2 sum_n sum_r (V_s(r, n) - target)^2   (slip rate at the receivers)
  J_w  = 1/2 sum_s |w(s) - target|^2            (stations, last step)
with respect to Dc, mu_s, mu_d and the initial shear traction, along seeded
smooth directions, against FP64 central differences of the FP64 oracle (on
its own setup), to 1e-3, with steps of about 1e-6 of each parameter (see
test_gpu_adjoint_fault.py for why the steps are that small).
"""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NT = 40
PARAMS = ["mu_s", "mu_d", "Dc", "C0", "Tx", "Ty", "Tz"]


def _smooth(shape, seed, passes=3):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape)
    for _ in range(passes):
        for ax in range(a.ndim):
            a = (np.roll(a, 1, ax) + 2 * a + np.roll(a, -1, ax)) / 4
    return a / np.abs(a).max()


def _scenario():
    from cgfd import configs
    sc = configs.tpv_like(ni=160, nj=240, nk=100, dh=100.0, pml=12, fault_len_km=12.0, fault_dep_km=6.0,
                          nstations=32, seed=4004, name="diffgrid")
    rng = np.random.default_rng(4004)
    sc.mu = (sc.mu * (1.0 + 0.03 * rng.standard_normal(sc.mu.shape))).astype(np.float32)
    f = sc.faults[0]
    i0 = f.x_index
    j1, j2, k1, k2 = f.grid
    jc, kc = (j1 + j2) // 2, sc.nk - 1 - 20  # the patch centre (2 km deep)
    near = [[i0 + dx, jc + dj, kc + dk] for dx, dj, dk in ((-2, 0, 0), (2, 3, 2), (-4, -5, 4), (5, 6, -3),
                                                            (-1, 10, 8), (3, -12, 10), (-6, 2, -6), (1, -3, 14))]
    sc.stations = np.concatenate([sc.stations, np.array(near)]).astype(np.int32)
    frec = np.array([[0, jc + dj + 3, kc + dk + 3] for dj in range(-14, 15, 4) for dk in range(-14, 15, 4)], np.int32)
    return sc, frec


def _fault_dir(sc, seed):
    j1, j2, k1, k2 = sc.faults[0].grid
    d = np.zeros((sc.nk, sc.nj))
    d[k1:k2 + 1, j1:j2 + 1] = _smooth((k2 - k1 + 1, j2 - j1 + 1), seed)
    return d


def _gpu_records(sc, frec, tape=False):
    from cgfd import model
    st = model.Setup(sc)
    core = model.Core(st)
    core.set_stations(sc.stations + 3, NT)
    core.set_fault_receivers(frec, NT)
    if tape:
        core.adjoint_forward(NT)
    core.run(NT)
    return st, core, core.read_stations(), core.read_fault_receivers()


def _fp64_records(sc, frec):
    import oracle
    orc = oracle.Oracle.from_scenario(sc, double=True)
    vs = []
    for _ in range(NT - 1):
        f = orc.read_fault(0)
        vs.append([f[3, k, j] for (_, j, k) in frec])
        orc.run(1)
    f = orc.read_fault(0)
    vs.append([f[3, k, j] for (_, j, k) in frec])
    x = orc.read_wavefield()
    last = np.array([x[:, k, j, i] for (i, j, k) in sc.stations + 3])
    orc.close()
    return np.array(vs).T, last


def _j_fp64(sc, frec, name, d, eps, vs_t, w_t):
    """(J_slip_rate, J_stations) of the FP64 oracle with `name` moved by
    eps d (runs in a worker process)."""
    s2 = copy.deepcopy(sc)
    s2.faults[0].fields[name] = (s2.faults[0].fields[name].astype(np.float64) + eps * d).astype(np.float32)
    vs, last = _fp64_records(s2, frec)
    return np.array([0.5 * np.sum((vs[:, 1:] - vs_t[:, 1:]) ** 2), 0.5 * np.sum((last - w_t) ** 2)])


def test_friction_gradients_differentiation_grid(built_libs):
    sc, frec = _scenario()
    sct = copy.deepcopy(sc)
    ft = sct.faults[0].fields
    ft["Dc"] = (ft["Dc"] * (1.0 + 0.1 * _fault_dir(sc, 4005))).astype(np.float32)
    ft["Ty"] = (ft["Ty"] + 0.1 * 70e6 * 0.05 * _fault_dir(sc, 4006)).astype(np.float32)
    stt, ct, tr_t, fr_t = _gpu_records(sct, frec)
    ct.close()
    stt.close()
    vs_t, w_t = fr_t[:, :, 3].astype(np.float64), tr_t[:, -1].astype(np.float64)

    st, core, tr, fr = _gpu_records(sc, frec, tape=True)
    print(f"[diffgrid] peak slip rate at the receivers {np.abs(fr[:, :, 3]).max():.3f} m/s", flush=True)
    assert np.abs(fr[:, :, 3]).max() > 0.05
    r_vs = np.zeros_like(fr)
    r_vs[:, :, 3] = fr[:, :, 3] - vs_t
    _, g_vs = core.adjoint_fault(np.zeros_like(tr), r_vs)
    r_w = np.zeros_like(tr)
    r_w[:, -1] = tr[:, -1] - w_t
    _, g_w = core.adjoint_fault(r_w)
    from test_gpu_adjoint_fault import _smooth_branch_mask
    mask = _smooth_branch_mask(sc, NT, core.read_fault(0))
    print(f"[diffgrid] directions restricted to {int(mask.sum())} smooth-branch nodes", flush=True)
    core.close()
    st.close()

    import fdpool
    params = (("Dc", 4e-7), ("Ty", 70.0), ("mu_s", 1e-6), ("mu_d", 1e-6))
    jobs = []
    for q, (name, eps) in enumerate(params):
        d = _fault_dir(sc, 4100 + q) * mask
        jobs += [(sc, frec, name, d, eps, vs_t, w_t), (sc, frec, name, d, -eps, vs_t, w_t)]
    print(f"[diffgrid] {len(jobs)} FP64 oracle runs of {NT} steps ...", flush=True)
    Js = fdpool.run_all(_j_fp64, jobs, workers=4)
    worst = 0.0
    for q, (name, eps) in enumerate(params):
        d = _fault_dir(sc, 4100 + q) * mask
        base = sc.faults[0].fields[name].astype(np.float64)
        step = (base + eps * d).astype(np.float32).astype(np.float64) - (base - eps * d).astype(np.float32).astype(np.float64)
        fd = Js[2 * q] - Js[2 * q + 1]
        for t, (lab, g) in enumerate((("slip rate", g_vs), ("last-step stations", g_w))):
            ad = float(np.sum(g[0, PARAMS.index(name)].astype(np.float64) * step))
            rel = abs(ad - fd[t]) / max(abs(fd[t]), 1e-300)
            worst = max(worst, rel)
            print(f"[diffgrid] {lab} d{name}: adjoint {ad:.8e}, FP64 central difference {fd[t]:.8e}, rel {rel:.2e}",
                  flush=True)
    assert worst <= 1e-3
