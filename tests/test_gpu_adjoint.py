"""Discrete adjoint of the wave update (interior velocity-stress stencil,
CFS-PML, free surface, RK4; cores without a fault) through the C ABI: no
boundary layer, CFS-PML on all six faces (edges and corners included), and
a free surface on top of CFS-PML on the other five faces.

* Linear identity: the station traces are a linear map T of the initial
  wavefield, so for any residual r the adjoint g = T^T r satisfies
  <g, w0> = sum r . T(w0) (both sides in FP64 from FP32 data).
* Least-squares misfit J = 1/2 sum (T(w0) - target)^2: the adjoint gradient
  against FP64 central differences of J computed by the FP64 oracle along
  seeded random directions (J is quadratic, so the central difference has no
  truncation error); agreement within 1e-3 relative.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NT = 24


CASES = [(0, False), (5, False), (5, True)]  # (PML layers, free surface)


def _scenario(pml=0, free=False):
    from cgfd import configs
    sc = configs.small(imethod=2, curvilinear=free)
    sc.faults = []
    sc.pml_layers = [[pml, pml], [pml, pml], [pml, 0 if free else pml]]
    sc.free_top = free
    rng = np.random.default_rng(41)
    extra = np.stack([rng.integers(2, sc.ni - 2, 6), rng.integers(2, sc.nj - 2, 6), rng.integers(2, sc.nk - 2, 6)], 1)
    sc.stations = np.concatenate([sc.stations, extra]).astype(np.int32)
    return sc


def _smooth_field(sc, seed, amp=1.0):
    """Random wavefield on the physical points (ghosts zero), smoothed so a
    short run stays well resolved."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = sc.shape_ext
    w = np.zeros((9, nz, ny, nx), np.float32)
    core = rng.standard_normal((9, sc.nk, sc.nj, sc.ni))
    for ax in (1, 2, 3):
        core = (np.roll(core, 1, ax) + 2 * core + np.roll(core, -1, ax)) / 4
    scale = np.array([1.0] * 3 + [1e6] * 6)[:, None, None, None]
    w[:, 3:-3, 3:-3, 3:-3] = (amp * scale * core).astype(np.float32)
    return w


def _traces_gpu(st, sc, w0):
    from cgfd import model
    core = model.Core(st)
    core.set_stations(sc.stations + 3, NT)
    core.write_wavefield(w0)
    core.run(NT)
    tr = core.read_stations()
    return core, tr


@pytest.mark.parametrize("pml,free", CASES)
def test_adjoint_linear_identity(built_libs, pml, free):
    from cgfd import model
    sc = _scenario(pml, free)
    st = model.Setup(sc)
    w0 = _smooth_field(sc, 5)
    core, tr = _traces_gpu(st, sc, w0)
    rng = np.random.default_rng(9)
    r = rng.standard_normal(tr.shape).astype(np.float32)
    r[:, :, 3:] *= 1e-6  # stress residuals on the stress scale
    g = core.adjoint_initial(r)
    lhs = float(np.sum(g.astype(np.float64) * w0.astype(np.float64)))
    rhs = float(np.sum(r.astype(np.float64) * tr.astype(np.float64)))
    print(f"[adjoint pml={pml} free={free}] <g, w0> = {lhs:.9e}, sum r.T(w0) = {rhs:.9e}", flush=True)
    assert np.isfinite(g).all()
    assert abs(lhs - rhs) <= 1e-4 * abs(rhs)
    # ghosts carry no gradient
    assert not g[:, :3].any() and not g[:, -3:].any()
    core.close()
    st.close()


@pytest.mark.parametrize("pml,free", CASES)
def test_adjoint_least_squares_vs_fp64_central_differences(built_libs, pml, free):
    from cgfd import model
    import oracle
    sc = _scenario(pml, free)
    st = model.Setup(sc)
    w0 = _smooth_field(sc, 6)
    target = _smooth_field(sc, 7, amp=0.5)
    _, tr_target = _traces_gpu(st, sc, target)
    core, tr = _traces_gpu(st, sc, w0)
    g = core.adjoint_initial(tr - tr_target).astype(np.float64)
    ijk = sc.stations + 3

    def j_fp64(w):
        orc = oracle.Oracle(st.desc, sc.shape_ext, 0, double=True)
        orc.write_wavefield(w)
        rec = []
        for _ in range(NT):
            x = orc.read_wavefield()
            rec.append([x[:, k, j, i] for (i, j, k) in ijk])
            orc.run(1)
        orc.close()
        rec = np.transpose(np.array(rec, np.float64), (1, 0, 2))
        return 0.5 * float(np.sum((rec - tr_target.astype(np.float64)) ** 2))

    worst = 0.0
    for d_seed in (101, 102, 103):
        d = _smooth_field(sc, d_seed, amp=0.1).astype(np.float64)
        eps = 1.0
        fd = (j_fp64(w0 + eps * d) - j_fp64(w0 - eps * d)) / (2 * eps)
        ad = float(np.sum(g * d))
        rel = abs(ad - fd) / abs(fd)
        worst = max(worst, rel)
        print(f"[adjoint pml={pml} free={free}] direction {d_seed}: adjoint {ad:.8e}, FP64 central difference {fd:.8e}, rel {rel:.2e}",
              flush=True)
    assert worst <= 1e-3
    core.close()
    st.close()


def test_adjoint_identity_differentiation_grid(built_libs):
    """The same linear identity on the differentiation grid (160 x 240 x 100
    nodes at 100 m, PML 12, free surface, 32 surface stations), 100 steps."""
    from cgfd import configs, model
    sc = configs.tpv_like(ni=160, nj=240, nk=100, dh=100.0, pml=12, fault_len_km=12.0, fault_dep_km=6.0,
                          nstations=32, seed=4004, name="config4-wave")
    sc.faults = []
    st = model.Setup(sc)
    nt = 100
    core = model.Core(st)
    core.set_stations(sc.stations + 3, nt)
    w0 = _smooth_field(sc, 11)
    core.write_wavefield(w0)
    core.run(nt)
    tr = core.read_stations()
    rng = np.random.default_rng(12)
    r = rng.standard_normal(tr.shape).astype(np.float32)
    r[:, :, 3:] *= 1e-6
    g = core.adjoint_initial(r)
    lhs = float(np.sum(g.astype(np.float64) * w0.astype(np.float64)))
    rhs = float(np.sum(r.astype(np.float64) * tr.astype(np.float64)))
    print(f"[adjoint differentiation grid] <g, w0> = {lhs:.9e}, sum r.T(w0) = {rhs:.9e}", flush=True)
    assert abs(lhs - rhs) <= 1e-4 * abs(rhs)
    core.close()
    st.close()
