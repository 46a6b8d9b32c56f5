"""GPU core vs the CPU oracle (FP32 restatement of the reference, on its own
independent FP64 setup of each scenario, oracle/setup_ref.py) on seeded
small configurations that exercise every path of the core: interior stencil,
CFS-PML on five or six faces, free-surface traction imaging and vlow,
the split-node fault with slip-weakening friction (both dynamic methods),
topography, two faults, and the fault2wave compatibility switch.

Tolerances (stated by the target): relative L2 <= 1e-3 for wavefield
snapshots, final slip and peak slip rate; rupture time within one dt.
"""
import numpy as np
import pytest

from helpers import fault_phys, interior, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-3


def _run_pair(sc, nsteps, compat=0, double_too=False):
    from cgfd import model
    import oracle
    st = model.Setup(sc)
    core = model.Core(st, compat_flags=compat)
    orc = oracle.Oracle.from_scenario(sc, double=False, fix_q1=bool(compat & 1))
    core.run(nsteps)
    orc.run(nsteps)
    out = {"gpu": core, "ora": orc, "setup": st}
    if double_too:
        o64 = oracle.Oracle.from_scenario(sc, double=True, fix_q1=bool(compat & 1))
        o64.run(nsteps)
        out["o64"] = o64
    return out


def _check(res, sc, label):
    core, orc = res["gpu"], res["ora"]
    wg = interior(core.read_wavefield())
    wo = interior(orc.read_wavefield())
    assert np.isfinite(wg).all(), f"{label}: non-finite GPU wavefield"
    names = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"]
    for c in range(9):
        r = rel_l2(wg[c], wo[c])
        assert r <= TOL, f"{label}: {names[c]} rel L2 {r:.3e} > {TOL}"
    for fid in range(len(sc.faults)):
        fg = fault_phys(core.read_fault(fid))
        fo = fault_phys(orc.read_fault(fid))
        r_slip = rel_l2(fg[6], fo[6])
        r_pv = rel_l2(fg[9], fo[9])
        assert r_slip <= TOL, f"{label}: fault {fid} slip rel L2 {r_slip:.3e}"
        assert r_pv <= TOL, f"{label}: fault {fid} peak slip rate rel L2 {r_pv:.3e}"
        rup_o = fo[10] > 0
        rup_g = fg[10] > 0
        assert (rup_o == rup_g).mean() >= 0.995, f"{label}: ruptured-node sets differ"
        both = rup_o & rup_g
        if both.any():
            dtt = np.abs(fg[10][both] - fo[10][both])
            assert dtt.max() <= sc.dt * 1.0001 + 1e-6, f"{label}: rupture time off by {dtt.max():.4g} s"
    return wg, wo


def test_small_flat_method2(built_libs):
    from cgfd import configs
    sc = configs.small(imethod=2)
    res = _run_pair(sc, 80)
    _check(res, sc, "flat m2")
    f = fault_phys(res["ora"].read_fault(0))
    assert (f[10] > 0).sum() > 50, "the test case must rupture"


def test_small_flat_method1(built_libs):
    from cgfd import configs
    sc = configs.small(imethod=1)
    res = _run_pair(sc, 80)
    _check(res, sc, "flat m1")


def test_small_curvilinear(built_libs):
    from cgfd import configs
    sc = configs.small(imethod=2, curvilinear=True)
    res = _run_pair(sc, 60)
    _check(res, sc, "curvilinear")


def test_no_free_surface_six_pml_faces(built_libs):
    from cgfd import configs
    sc = configs.small(imethod=2, free_top=False)
    res = _run_pair(sc, 60)
    _check(res, sc, "buried fault, 6 PML faces")


def test_two_faults(built_libs):
    from cgfd import configs
    sc = configs.small(ni=48, imethod=2, nfault=2)
    res = _run_pair(sc, 50)
    _check(res, sc, "two faults")


def test_fault2wave_swap_fix(built_libs):
    from cgfd import configs
    sc = configs.small(imethod=1)
    res = _run_pair(sc, 40, compat=1)
    _check(res, sc, "compat fix q1")


def test_pg_and_stations(built_libs):
    from cgfd import configs, model
    import oracle
    sc = configs.small(imethod=2)
    st = model.Setup(sc)
    core = model.Core(st)
    ijk = sc.stations + 3
    core.set_stations(ijk, 64)
    orc = oracle.Oracle.from_scenario(sc)
    rec_o = []
    for _ in range(40):
        w = orc.read_wavefield()
        rec_o.append([w[:, k, j, i] for (i, j, k) in ijk])
        orc.run(1)
    core.run(40)
    rec_g = core.read_stations()
    rec_o = np.transpose(np.array(rec_o), (1, 0, 2))
    assert rec_g.shape == rec_o.shape
    assert rel_l2(rec_g, rec_o) <= TOL
    pg_g = core.read_pg()[:, 3:-3, 3:-3]
    pg_o = orc.read_pg()[:, 3:-3, 3:-3]
    for c in range(15):
        assert rel_l2(pg_g[c], pg_o[c]) <= TOL, f"PG component {c}"


def test_fp64_oracle_error_statement(built_libs):
    """The FP32 GPU result vs the FP64 oracle: reported, bounded loosely."""
    from cgfd import configs
    sc = configs.small(imethod=2)
    res = _run_pair(sc, 60, double_too=True)
    wg = interior(res["gpu"].read_wavefield())
    w64 = interior(res["o64"].read_wavefield())
    w32 = interior(res["ora"].read_wavefield())
    e_gpu = rel_l2(wg, w64)
    e_ora = rel_l2(w32, w64)
    print(f"rel L2 vs FP64: gpu {e_gpu:.3e}, fp32 oracle {e_ora:.3e}")
    assert e_gpu <= TOL


@pytest.mark.parametrize("variant", [1, 20, 23, 24, 33])
def test_volume_variants_match_oracle(built_libs, variant):
    """Every volume-stencil implementation the autotuner chooses from."""
    from cgfd import configs, model
    import oracle
    sc = configs.small(ni=70, nj=52, nk=40, imethod=2, curvilinear=True)
    st = model.Setup(sc)
    core = model.Core(st, variant=variant)
    orc = oracle.Oracle.from_scenario(sc)
    core.run(24)
    orc.run(24)
    wg = interior(core.read_wavefield())
    wo = interior(orc.read_wavefield())
    assert rel_l2(wg, wo) <= TOL


def test_graph_replay_matches_eager(built_libs):
    """Eight-step graphs replayed several times, with eager steps and
    synchronisations in between, against direct launches: bit-identical
    wavefield, fault outputs and stations after every replay."""
    from cgfd import configs, model
    sc = configs.small(imethod=2)
    st = model.Setup(sc)
    ijk = sc.stations + 3
    cg = model.Core(st)
    ce = model.Core(st)
    ce.set_graph(False)
    for c in (cg, ce):
        c.set_stations(ijk, 64)
    for rep, n in enumerate((3, 13, 8, 16, 1)):
        cg.run(n)
        ce.run(n)
        cg.sync()
        ce.sync()
        assert cg.step_index == ce.step_index
        assert np.array_equal(cg.read_wavefield(), ce.read_wavefield()), f"wavefield differs after block {rep}"
        assert np.array_equal(cg.read_fault(0), ce.read_fault(0)), f"fault differs after block {rep}"
    assert np.array_equal(cg.read_stations(), ce.read_stations())


@pytest.mark.parametrize("variant", [20, 1, 23, 24])
def test_stage_fork_matches_serial(built_libs, variant):
    """The stage fork (interior blocks, CFS-PML blocks + top layers, fault
    kernels on three streams) against every kernel on one stream: the
    kernels write disjoint points, so the results are bit-identical, eager
    and replayed from graphs."""
    from cgfd import configs, model
    sc = configs.small(imethod=2)
    st = model.Setup(sc)
    cf = model.Core(st, variant=variant)
    cf.set_overlap(True)
    cs = model.Core(st, variant=variant)
    cs.set_overlap(False)
    cs.set_graph(False)
    for n in (5, 16, 3):
        cf.run(n)
        cs.run(n)
        assert np.array_equal(cf.read_wavefield(), cs.read_wavefield())
        assert np.array_equal(cf.read_fault(0), cs.read_fault(0))
    assert np.array_equal(cf.read_pg(), cs.read_pg())
    cf.close()
    cs.close()
    st.close()


def test_state_save_load_restarts_bit_identically(built_libs):
    """Step-boundary state saved to host memory and to device memory, the run
    continued, the state loaded back (also into a second core) and the same
    steps run again: bit-identical wavefield, fault outputs, PG and step
    index. The restored run crosses a graph block boundary."""
    import ctypes as C
    from cgfd import configs, model
    sc = configs.small(imethod=2)
    st = model.Setup(sc)
    core = model.Core(st)
    core.run(13)
    host = core.save_state()
    # device buffer from the HIP runtime the core library is linked against
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    dptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(dptr), core.state_bytes()) == 0
    core.save_state(dptr.value)
    core.run(11)
    ref = (core.read_wavefield(), core.read_fault(0), core.read_pg(), core.step_index)
    for how in ("host", "device"):
        if how == "host":
            core.load_state(host)
        else:
            core.load_state(dev_ptr=dptr.value)
        assert core.step_index == 13
        core.run(11)
        got = (core.read_wavefield(), core.read_fault(0), core.read_pg(), core.step_index)
        assert got[3] == ref[3]
        for a, b in zip(got[:3], ref[:3]):
            assert np.array_equal(a, b), f"restart from {how} memory differs"
    other = model.Core(st)
    other.load_state(host)
    other.run(11)
    assert np.array_equal(other.read_wavefield(), ref[0])
    assert np.array_equal(other.read_fault(0), ref[1])
    other.close()
    core.close()
    assert hip.hipFree(dptr) == 0
    st.close()


@pytest.mark.parametrize("imethod", [1, 2])
def test_locked_fault_transmits_without_slip_gpu(built_libs, imethod):
    """The GPU fault kernels on their own: a locked fault (static friction
    1000) crossed by a velocity pulse keeps the slip rate at FP32 rounding
    level, as the trial traction enforces velocity continuity."""
    from cgfd import configs, model
    sc = configs.small(imethod=imethod)
    sc.faults[0].fields["mu_s"][:] = 1000.0
    i0 = sc.faults[0].x_index + 3
    st = model.Setup(sc)
    core = model.Core(st)
    nz, ny, nx = sc.shape_ext
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    g = np.exp(-(((ii - (i0 - 8)) / 2.5) ** 2 + ((jj - ny // 2) / 2.5) ** 2 + ((kk - nz // 2) / 2.5) ** 2))
    w0 = np.zeros((9, nz, ny, nx), np.float32)
    w0[0], w0[1], w0[2] = g, 0.5 * g, -0.3 * g
    w0[:, :3] = w0[:, -3:] = 0
    w0[:, :, :3] = w0[:, :, -3:] = 0
    w0[:, :, :, :3] = w0[:, :, :, -3:] = 0
    core.write_wavefield(w0)
    vmax = vsmax = 0.0
    for _ in range(12):
        core.run(5)
        w = core.read_wavefield()
        vmax = max(vmax, float(np.abs(w[:3, :, :, i0 - 3:i0 + 4]).max()))
        vsmax = max(vsmax, float(np.abs(core.read_fault(0)[3]).max()))
    assert vmax > 0.05
    assert vsmax <= 1e-5 * vmax, f"slip rate {vsmax:.3e} on a locked fault (particle velocity {vmax:.3e})"
    core.close()
    st.close()


@pytest.mark.parametrize("variant", [20, 1, 23, 24])
def test_ragged_grid_matches_oracle(built_libs, variant):
    """A grid whose sizes are multiples of none of the tile sizes (37 x 45 x
    27: partial 64-wide x tiles, a partial last row of 4-row y tiles, z
    chunks that do not divide the planes) with a fault, PML and a free
    surface: every point is computed once, against the FP32 oracle."""
    from cgfd import configs, model
    import oracle
    sc = configs.small(ni=37, nj=45, nk=27, pml=5, imethod=2)
    st = model.Setup(sc)
    core = model.Core(st, variant=variant)
    core.set_zchunk(4)
    orc = oracle.Oracle.from_scenario(sc)
    core.run(30)
    orc.run(30)
    wg = interior(core.read_wavefield())
    wo = interior(orc.read_wavefield())
    assert rel_l2(wg, wo) <= TOL
    fg, fo = fault_phys(core.read_fault(0)), fault_phys(orc.read_fault(0))
    assert rel_l2(fg[6], fo[6]) <= TOL
    core.close()
    orc.close()
    st.close()


def test_slices_and_snapshots_match_the_wavefield(built_libs):
    """Device-packed slices and strided snapshots (io_slice_nc_put /
    io_snap_nc_put) against the full wavefield read back after a short
    rupture: bit-identical velocity and stress; strain against the isotropic
    compliance of io_snap_stress_to_strain_eliso (io_funcs.cu:1585-1651)
    evaluated in FP64 from the same stress and media."""
    from cgfd import configs, model
    sc = configs.small(imethod=2)
    st = model.Setup(sc)
    core = model.Core(st)
    core.run(25)
    w = core.read_wavefield()
    i0 = sc.faults[0].x_index + 3
    assert np.array_equal(core.read_slice(0, i0 - 1), w[:, 3:-3, 3:-3, i0 - 1])
    assert np.array_equal(core.read_slice(1, 20), w[:, 3:-3, 20, 3:-3])
    assert np.array_equal(core.read_slice(2, sc.nk + 2), w[:, sc.nk + 2, 3:-3, 3:-3])
    box = (4, 9, 3, 5, 7, 5, 6, 8, 2)
    snap = core.read_snapshot(box, 7)
    sl = (slice(box[6], box[6] + box[7] * box[8], box[8]), slice(box[3], box[3] + box[4] * box[5], box[5]),
          slice(box[0], box[0] + box[1] * box[2], box[2]))
    wb = w[(slice(None),) + sl]
    assert np.array_equal(snap[:9], wb)
    nz, ny, nx = sc.shape_ext
    pad = lambda a: np.pad(a, 3, mode="edge")
    lam, mu = pad(sc.lam.astype(np.float64))[sl], pad(sc.mu.astype(np.float64))[sl]
    e1 = (lam + mu) / (mu * (3 * lam + 2 * mu))
    e2 = -lam / (2 * mu * (3 * lam + 2 * mu))
    t = wb[3:].astype(np.float64)
    e0 = e2 * (t[0] + t[1] + t[2])
    ref = np.stack([e0 - (e2 - e1) * t[0], e0 - (e2 - e1) * t[1], e0 - (e2 - e1) * t[2],
                    0.5 / mu * t[3], 0.5 / mu * t[4], 0.5 / mu * t[5]])
    err = np.abs(snap[9:] - ref).max() / np.abs(ref).max()
    print(f"[snapshot] strain max rel err {err:.2e}", flush=True)
    assert err <= 1e-6
    assert np.array_equal(core.read_snapshot(box, 4), snap[9:])
    core.close()
    st.close()


@pytest.mark.parametrize("variant", [23, 24, 33])
def test_box_split_ragged_grid_matches_oracle(built_libs, variant):
    """The box-split schedule where its interior box is not empty and no
    size is a multiple of a tile: 133 x 61 x 35 nodes with PML 5 and a free
    surface give an interior box of columns 32..127 (three 32-wide or one
    and a half 64-wide tiles), partial 4- or 8-row tiles and six-sided
    bands around it, with the fault plane inside the box."""
    from cgfd import configs, model
    import oracle
    sc = configs.small(ni=133, nj=61, nk=35, pml=5, imethod=2, curvilinear=True)
    st = model.Setup(sc)
    core = model.Core(st, variant=variant)
    core.set_zchunk(4)
    orc = oracle.Oracle.from_scenario(sc)
    core.run(24)
    orc.run(24)
    r = rel_l2(interior(core.read_wavefield()), interior(orc.read_wavefield()))
    rs = rel_l2(fault_phys(core.read_fault(0))[6], fault_phys(orc.read_fault(0))[6])
    print(f"[box split {variant}] wavefield {r:.2e}, slip {rs:.2e}", flush=True)
    assert r <= TOL and rs <= TOL
    core.close()
    orc.close()
    st.close()


@pytest.mark.parametrize("variant", [23, 24, 33])
def test_vertical_column_form_matches_the_nine_term_form(built_libs, variant):
    """On a grid of vertical columns under topography (xi_y, xi_z, eta_x and
    eta_z hold only the metric's rounding, about 1e-15 of xi_x and eta_y)
    the box split reads five metric terms instead of nine; after 30 steps
    of rupture its results agree with the nine-term form to 1e-6 (they
    differ in at most an ulp per operation, at few points), and on an
    affine grid the core does not take it."""
    from cgfd import configs, model
    import pins
    sc = configs.small(ni=133, nj=61, nk=35, pml=5, imethod=2, curvilinear=True)
    st = model.Setup(sc)
    a = model.Core(st, variant=variant)
    b = model.Core(st, variant=variant)
    b.set_vertical_columns(False)
    assert a.vertical_columns() == (True, True) and b.vertical_columns() == (True, False)
    a.run(30)
    b.run(30)
    wa, wb = a.read_wavefield(), b.read_wavefield()
    fa, fb = a.read_fault(0), b.read_fault(0)
    assert np.abs(wa[:3]).max() > 0
    same = float(np.mean(wa == wb))
    rw = max(rel_l2(wa[c], wb[c]) for c in range(9))
    rf = rel_l2(fault_phys(fa)[6], fault_phys(fb)[6])
    print(f"[vertical columns {variant}] bitwise equal {100 * same:.4f}% of wavefield values, "
          f"worst component rel L2 {rw:.2e}, slip rel L2 {rf:.2e}", flush=True)
    assert rw <= 1e-6 and rf <= 1e-6
    for c in (a, b):
        c.close()
    st.close()
    A = pins.affine_matrix(100.0)
    sa = pins.affine_scenario(40, 36, 30, A, 1e-3)
    sta = model.Setup(sa)
    c = model.Core(sta, variant=variant)
    assert c.vertical_columns() == (False, False)
    c.close()
    sta.close()
