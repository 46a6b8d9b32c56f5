"""Discrete adjoint of the time loop with a split-node slip-weakening fault,
through the C ABI (cgfd_adj_forward / cgfd_adj_backward_fault).

* Locked fault (static friction 1000, compressive normal traction): the
  friction never yields, so the station traces are a linear map T of the
  initial wavefield through every fault path (wave2fault, traction columns,
  trial, fault2wave, fault columns, split nodes, hT1 recurrence, free-surface
  rows). The adjoint g = T^T r must satisfy <g, w0> = sum r . T(w0), for both
  split-node methods. Stations sit next to the plane and on it.
* Rupturing fault, 40 and 200 steps: J_tr = 1/2 sum (T(theta) - target)^2 over the
  station traces and J_vs = 1/2 sum (V_s(theta) - V_s target)^2 over the slip
  rate at fault receivers, as functions of the friction parameters theta;
  the adjoint gradients along 20 seeded smooth directions each of mu_s, mu_d,
  Dc and the initial shear traction against FP64 central differences of J
  computed by the FP64 oracle, to 1e-3 (see the test for the bound at 200
  steps). The friction law is differentiated
  on the branches the FP32 forward run took, and the difference quotient is
  of the FP64 discrete map, which is piecewise smooth. The directions are
  therefore restricted to the nodes whose friction branch decisions stay at
  least 1e-4 (relative) from their switch points over the whole FP64 run
  (the oracle records those margins), where the FP32 and FP64 runs and the
  +-step runs all take the same branches. Steps of about 1e-6 of each
  parameter (a few FP32 units in the last place; the quotient uses the
  perturbation actually taken) and FP64 keep the quotient accurate to ~1e-9.
* A core with a fault and no tape is refused.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PARAMS = ["mu_s", "mu_d", "Dc", "C0", "Tx", "Ty", "Tz"]


def _scenario(imethod, locked, nfault=1, free_top=True):
    from cgfd import configs
    sc = configs.small(imethod=imethod, nfault=nfault, free_top=free_top)
    f = sc.faults[0]
    if locked:
        for ff in sc.faults:
            ff.fields["mu_s"][:] = 1000.0
    i0 = f.x_index
    j1, j2, k1, k2 = f.grid
    jc = (j1 + j2) // 2
    extra = [[i0 - 1, jc, sc.nk - 3], [i0 + 1, jc + 2, sc.nk - 5], [i0 - 3, jc - 3, sc.nk - 2],
             [i0 + 2, jc + 1, sc.nk - 1], [i0, jc - 1, sc.nk - 4], [i0, jc + 3, sc.nk - 1],
             [i0 + 4, jc - 2, sc.nk - 6], [i0 - 5, jc + 4, sc.nk - 1]]
    sc.stations = np.concatenate([sc.stations, np.array(extra)]).astype(np.int32)
    return sc


def _smooth_field(sc, seed, amp):
    rng = np.random.default_rng(seed)
    nz, ny, nx = sc.shape_ext
    w = np.zeros((9, nz, ny, nx), np.float32)
    core = rng.standard_normal((9, sc.nk, sc.nj, sc.ni))
    for ax in (1, 2, 3):
        core = (np.roll(core, 1, ax) + 2 * core + np.roll(core, -1, ax)) / 4
    scale = np.array([1.0] * 3 + [1e6] * 6)[:, None, None, None]
    w[:, 3:-3, 3:-3, 3:-3] = (amp * scale * core).astype(np.float32)
    return w


@pytest.mark.parametrize("imethod,nfault,free_top", [(1, 1, True), (2, 1, True), (2, 2, True), (1, 2, True),
                                                     (2, 1, False), (1, 1, False)])
def test_fault_adjoint_linear_identity_locked(built_libs, imethod, nfault, free_top):
    from cgfd import model
    nt = 30
    sc = _scenario(imethod, locked=True, nfault=nfault, free_top=free_top)
    st = model.Setup(sc)
    core = model.Core(st)
    core.set_stations(sc.stations + 3, nt)
    core.adjoint_forward(nt)
    w0 = _smooth_field(sc, 5, 1e-3)
    core.write_wavefield(w0)
    core.run(nt)
    tr = core.read_stations()
    rng = np.random.default_rng(9)
    r = rng.standard_normal(tr.shape).astype(np.float32)
    r[:, :, 3:] *= 1e-6
    g, gf = core.adjoint_fault(r)
    lhs = float(np.sum(g.astype(np.float64) * w0.astype(np.float64)))
    rhs = float(np.sum(r.astype(np.float64) * tr.astype(np.float64)))
    print(f"[fault adjoint locked, method {imethod}, {nfault} fault(s), free top {free_top}] <g, w0> = {lhs:.9e}, "
          f"sum r.T(w0) = {rhs:.9e}, "
          f"rel {abs(lhs - rhs) / abs(rhs):.2e}", flush=True)
    assert np.isfinite(g).all() and np.isfinite(gf).all()
    assert abs(lhs - rhs) <= 1e-4 * abs(rhs)
    core.close()
    st.close()


def _traces_fp64(sc, nt, frec):
    """Station traces and fault-receiver slip rates of the FP64 oracle (on
    its own setup of the scenario)."""
    import oracle
    orc = oracle.Oracle.from_scenario(sc, double=True)
    ijk = sc.stations + 3
    rec, vs = [], []
    for _ in range(nt):
        x = orc.read_wavefield()
        rec.append([x[:, k, j, i] for (i, j, k) in ijk])
        f = orc.read_fault(0)
        vs.append([f[3, k, j] for (_, j, k) in frec])
        orc.run(1)
    orc.close()
    return np.transpose(np.array(rec, np.float64), (1, 0, 2)), np.array(vs, np.float64).T


def _j_fp64(sc, name, d, eps, nt, frec, target, vs_target):
    """(J_traces, J_slip_rate) of the FP64 oracle with `name` moved by eps d
    (runs in a worker process)."""
    import copy
    s2 = copy.deepcopy(sc)
    s2.faults[0].fields[name] = (s2.faults[0].fields[name].astype(np.float64) + eps * d).astype(np.float32)
    rec, vs = _traces_fp64(s2, nt, frec)
    return np.array([0.5 * float(np.sum((rec - target) ** 2)), 0.5 * float(np.sum((vs[:, 1:] - vs_target[:, 1:]) ** 2))])


def _smooth_branch_mask(sc, nt, fr_out_gpu, rel=1e-4):
    """Fault nodes at which every branch decision of the friction law
    (trial_slipweakening.cu:320-364: shear traction against strength, slip
    against Dc, normal traction against 0, the 0.1 guard) stays more than rel
    (relative) away from its switch point at every trial evaluation of the
    FP64 run, and whose rupture onset step is the one the FP32 GPU run took.
    FP32 and FP64 trajectories differ by far less than rel here, so at these
    nodes the FP32 branches the adjoint follows are those of the FP64 map the
    difference quotient sees, and the +-step runs stay on them too; at the
    other nodes a kink can fall between the two runs or inside the step."""
    import oracle
    orc = oracle.Oracle.from_scenario(sc, double=True)
    orc.run(nt)
    margin = orc.read_kink_margin(0)[:, 3:-3, 3:-3].min(axis=0)
    t0_64 = orc.read_fault(0)[10, 3:-3, 3:-3]
    orc.close()
    t0_32 = fr_out_gpu[10, 3:-3, 3:-3].astype(np.float64)
    same_onset = np.abs(t0_64 - t0_32) < 0.5 * sc.dt
    return ((margin > rel) & same_onset).astype(np.float64)


def _fault_receivers(sc):
    """(fault id, ghost-inclusive j, k) around the nucleation patch"""
    j1, j2, k1, k2 = sc.faults[0].grid
    jc, kc = (j1 + j2) // 2, sc.nk - 1 - 9
    return np.array([[0, jc + dj + 3, kc + dk + 3] for dj in (-6, -3, 0, 2, 5) for dk in (-4, 0, 3)], np.int32)


def _direction(sc, name, seed):
    """Smooth seeded field on the fault's own grid (zero outside it)."""
    rng = np.random.default_rng(seed)
    j1, j2, k1, k2 = sc.faults[0].grid
    d = np.zeros((sc.nk, sc.nj))
    core = rng.standard_normal((k2 - k1 + 1, j2 - j1 + 1))
    for ax in (0, 1):
        core = (np.roll(core, 1, ax) + 2 * core + np.roll(core, -1, ax)) / 4
    d[k1:k2 + 1, j1:j2 + 1] = core
    return d


# parameter, step size of the difference quotient (units of the parameter):
# about 1e-6 of mu (0.5-0.7), Dc (0.1 m) and the shear traction (~70 MPa)
STEPS = [("mu_s", 1e-6), ("mu_d", 1e-6), ("Dc", 1e-7), ("Ty", 70.0)]
NDIR = 20


@pytest.mark.parametrize("imethod,nt", [(2, 40), (1, 40), (2, 200), (1, 200)])
def test_fault_adjoint_parameter_gradient_vs_fp64(built_libs, imethod, nt):
    """For each parameter and functional, the 20 directional derivatives as
    a vector (a projection of the gradient) to 1e-3 in relative L2, and
    their median per-direction relative error to 1e-3; at 40 steps also
    every direction whose difference quotient is within 10x of the largest
    to 1e-3 of its own size (directions far smaller than that are dominated
    by cancellation in the sum over nodes and are reported, not bounded).
    Over 200 steps the slip pulses arrest and re-stick nodes, so nearly
    every node comes close to a stick/slip switch at some step, where the
    gradient is sensitive to the 1e-6-level difference between the FP32
    trajectory the adjoint follows and the FP64 one the difference quotient
    sees; single directions then drift to a few 1e-3 of their own size while
    the vector stays within 1e-3."""
    from cgfd import model
    import copy
    sc = _scenario(imethod, locked=False)
    frec = _fault_receivers(sc)
    st = model.Setup(sc)
    core = model.Core(st)
    core.set_stations(sc.stations + 3, nt)
    core.set_fault_receivers(frec, nt)
    core.adjoint_forward(nt)
    core.run(nt)
    tr = core.read_stations()
    fr = core.read_fault_receivers()
    # target: the traces of a seeded perturbation of Dc and the shear traction
    sct = copy.deepcopy(sc)
    sct.faults[0].fields["Dc"] += (0.05 * sc.faults[0].fields["Dc"] * _direction(sc, "Dc", 77)).astype(np.float32)
    sct.faults[0].fields["Ty"] += (2e5 * _direction(sc, "Ty", 78)).astype(np.float32)
    stt = model.Setup(sct)
    ct = model.Core(stt)
    ct.set_stations(sct.stations + 3, nt)
    ct.set_fault_receivers(frec, nt)
    ct.run(nt)
    target = ct.read_stations().astype(np.float64)
    vs_target = ct.read_fault_receivers()[:, :, 3].astype(np.float64)
    ct.close()
    stt.close()
    assert float(np.abs(fr[:, :, 3]).max()) > 0.05, "the receivers should see the rupture"
    r = (tr - target).astype(np.float32)
    rf = np.zeros_like(fr)
    rf[:, :, 3] = fr[:, :, 3] - vs_target
    _, gf_tr = core.adjoint_fault(r)
    _, gf_vs = core.adjoint_fault(np.zeros_like(r), rf)
    assert np.isfinite(gf_tr).all() and np.isfinite(gf_vs).all()

    import fdpool
    mask = _smooth_branch_mask(sc, nt, fr_out_gpu=core.read_fault(0))
    grid = _direction(sc, "Dc", 1) != 0
    print(f"[fault adjoint method {imethod}, {nt} steps] directions restricted to {int(mask[grid].sum())} of "
          f"{int(grid.sum())} fault-grid nodes (the others come within 1e-4 of a friction branch switch in the "
          f"FP64 run or start slipping at a different step in FP32 and FP64)", flush=True)
    jobs, dirs = [], []
    for q, (name, eps) in enumerate(STEPS):
        for k in range(NDIR):
            d = _direction(sc, name, 200 + 100 * q + k) * mask
            dirs.append((q, name, eps, k, d))
            for sgn in (1.0, -1.0):
                jobs.append((sc, name, d, sgn * eps, nt, frec, target, vs_target))
    print(f"[fault adjoint method {imethod}, {nt} steps] {len(jobs)} FP64 oracle runs ...", flush=True)
    Js = fdpool.run_all(_j_fp64, jobs)
    tag = f"[fault adjoint method {imethod}, {nt} steps]"
    worst = 0.0
    for q, (name, eps) in enumerate(STEPS):
        for t, (lab, gf) in enumerate((("traces", gf_tr), ("slip rate", gf_vs))):
            ad, fd = [], []
            for n, (q2, name2, eps2, k, d) in enumerate(dirs):
                if q2 != q:
                    continue
                # the step actually taken in FP32 parameters
                base = sc.faults[0].fields[name].astype(np.float64)
                plus = (base + eps * d).astype(np.float32).astype(np.float64)
                minus = (base - eps * d).astype(np.float32).astype(np.float64)
                ad.append(float(np.sum(gf[0, PARAMS.index(name)].astype(np.float64) * (plus - minus))))
                fd.append(Js[2 * n][t] - Js[2 * n + 1][t])
            ad, fd = np.array(ad), np.array(fd)
            rel = np.abs(ad - fd) / np.maximum(np.abs(fd), 1e-300)
            l2 = float(np.linalg.norm(ad - fd) / np.linalg.norm(fd))
            big = np.abs(fd) >= 0.1 * np.abs(fd).max()
            kw = int(np.argmax(np.where(big, rel, -1.0)))
            print(f"{tag} {lab} d{name}: {NDIR} directions, rel L2 over the directions {l2:.2e}, per-direction "
                  f"median {np.median(rel):.2e}, worst {rel.max():.2e} (at |difference| "
                  f"{np.abs(fd[np.argmax(rel)]) / np.abs(fd).max():.1e} of the largest), worst of the "
                  f"{int(big.sum())} directions within 10x of the largest {rel[kw]:.2e} (direction {kw}: adjoint "
                  f"{ad[kw]:.6e}, FP64 difference {fd[kw]:.6e})", flush=True)
            worst = max(worst, l2, float(np.median(rel)))
            if nt <= 40:
                worst = max(worst, rel[kw])
    assert worst <= 1e-3
    core.close()
    st.close()


def test_fault_adjoint_needs_tape(built_libs):
    from cgfd import configs, model
    sc = configs.small(imethod=2)
    st = model.Setup(sc)
    core = model.Core(st)
    core.set_stations(sc.stations + 3, 4)
    core.run(4)
    r = np.zeros((len(sc.stations), 4, 9), np.float32)
    nz, ny, nx = sc.shape_ext
    out = np.zeros((9, nz, ny, nx), np.float32)
    rc = core.lib.cgfd_adj_backward(core.handle, r.ctypes.data_as(C.c_void_p), 4, out.ctypes.data_as(C.c_void_p))
    assert rc == -4
    assert b"friction tape" in core.lib.cgfd_last_error()
    core.close()
    st.close()


def test_adjoint_cycles_reuse_buffers(built_libs):
    """An inversion loop restarts one core from its saved initial state and
    re-arms stations and tape every cycle: the second cycle's adjoint is as
    exact as the first, and re-arming does not grow device memory (the
    buffers are reused, not reallocated)."""
    import ctypes as C
    from cgfd import model
    nt = 20
    sc = _scenario(2, locked=True)
    st = model.Setup(sc)
    core = model.Core(st)
    s0 = core.save_state()
    for cycle in range(2):
        core.load_state(s0)
        core.set_stations(sc.stations + 3, nt)
        core.adjoint_forward(nt)
        w0 = _smooth_field(sc, 50 + cycle, 1e-3)
        core.write_wavefield(w0)
        core.run(nt)
        tr = core.read_stations()
        r = np.random.default_rng(60 + cycle).standard_normal(tr.shape).astype(np.float32)
        r[:, :, 3:] *= 1e-6
        g, _ = core.adjoint_fault(r)
        lhs = float(np.sum(g.astype(np.float64) * w0.astype(np.float64)))
        rhs = float(np.sum(r.astype(np.float64) * tr.astype(np.float64)))
        print(f"[adjoint cycle {cycle}] rel {abs(lhs - rhs) / abs(rhs):.2e}", flush=True)
        assert abs(lhs - rhs) <= 1e-4 * abs(rhs)
    hip = C.CDLL("libamdhip64.so.7")
    free0, free1, total = C.c_size_t(), C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free0), C.byref(total)) == 0
    for _ in range(20):
        core.set_stations(sc.stations + 3, nt)
        core.adjoint_forward(nt)
    assert hip.hipMemGetInfo(C.byref(free1), C.byref(total)) == 0
    grew = free0.value - free1.value
    print(f"[adjoint cycles] device memory taken by 20 re-arms: {grew / 2**20:.1f} MiB", flush=True)
    assert grew < 8 << 20
    core.close()
    st.close()
