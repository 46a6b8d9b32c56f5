"""The benchmark configurations through the C ABI at sizes a test can afford,
and the one-call drop-in and the autotuner.

* config 2 (the plain-method rupture on a curvilinear topographic grid) at a
  quarter of its linear size, with the core's default schedule, against the
  FP32 oracle (on its own setup) over 60 steps;
* config 3: the 32-scenario batch on the full 400 x 200 x 150 grid with run
  lengths scaled to 2% (20-60 steps, every scenario retiring at its own
  step), each scenario against its plain run; and one scenario of the same
  generator at half scale against the oracle;
* cgfd_drv_rk_curv_col_allstep with a real descriptor equals create / run /
  read bit for bit and agrees with the oracle;
* cgfd_core_autotune leaves the simulation state bit-identical, and the
  schedule it picks is recorded for the shape and passes parity.
"""
import copy
import ctypes as C

import numpy as np
import pytest

from helpers import fault_phys, interior, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-3


def _log(msg):
    print(msg, flush=True)


def _check_vs_oracle(wg, fg, orc, sc, label):
    wo = interior(orc.read_wavefield())
    r = rel_l2(interior(wg), wo)
    fo = fault_phys(orc.read_fault(0))
    f = fault_phys(fg)
    rs, rp = rel_l2(f[6], fo[6]), rel_l2(f[9], fo[9])
    rup_o, rup_g = fo[10] > 0, f[10] > 0
    both = rup_o & rup_g
    dtt = float(np.abs(f[10][both] - fo[10][both]).max()) if both.any() else 0.0
    _log(f"[{label}] wavefield {r:.2e}, slip {rs:.2e}, peak slip rate {rp:.2e}, ruptured nodes {int(rup_g.sum())}, "
         f"rupture time max diff {dtt:.4g} s")
    assert r <= TOL and rs <= TOL and rp <= TOL
    assert rup_g.sum() > 0 and (rup_o == rup_g).mean() >= 0.995
    assert dtt <= sc.dt * 1.0001 + 1e-6


def test_config2_quarter_scale_default_schedule_vs_oracle(built_libs):
    from cgfd import configs, model
    import oracle
    sc = configs.config2(scale=0.25, steps=60)
    _log(f"[config2/4] {sc.ni}x{sc.nj}x{sc.nk}, 60 steps")
    st = model.Setup(sc)
    core = model.Core(st)
    core.run(60)
    wg, fg = core.read_wavefield(), core.read_fault(0)
    core.close()
    st.close()
    orc = oracle.Oracle.from_scenario(sc)
    orc.run(60)
    _check_vs_oracle(wg, fg, orc, sc, "config2/4 default schedule")
    orc.close()


def test_config3_batch_of_32_vs_plain_runs(built_libs):
    from cgfd import configs, model
    base, fields, nts = configs.config3(nscen=32, steps_scale=0.02)
    _log(f"[config3] {base.ni}x{base.nj}x{base.nk}, 32 scenarios, run lengths {min(nts)}..{max(nts)}")
    assert len(set(nts)) > 8, "run lengths must be uneven"
    st = model.Setup(base)
    bat = model.Batch(st, [(None if s == 0 else [fields[s]], nts[s]) for s in range(32)])
    bat.run(max(nts) + 2)
    bat.sync()
    assert bat.active == 0
    worst = 0.0
    for s in range(32):
        assert bat.steps(s) == nts[s]
        sc = copy.copy(base)
        sc.faults = [copy.copy(base.faults[0])]
        sc.faults[0].fields = fields[s]
        st_s = model.Setup(sc)
        core = model.Core(st_s)
        core.run(nts[s])
        r = rel_l2(interior(bat.read_wavefield(s)), interior(core.read_wavefield()))
        rs = rel_l2(fault_phys(bat.read_fault(s))[6], fault_phys(core.read_fault(0))[6])
        worst = max(worst, r, rs)
        assert r <= 1e-6 and rs <= 1e-6, f"scenario {s}: wavefield {r:.2e}, slip {rs:.2e}"
        core.close()
        st_s.close()
    _log(f"[config3] batched vs plain, worst rel L2 over 32 scenarios: {worst:.2e}")
    bat.close()
    st.close()


def test_config3_half_scale_scenario_vs_oracle(built_libs):
    from cgfd import configs, model
    import oracle
    base, fields, nts = configs.config3(nscen=4, scale=0.5, steps_scale=0.02)
    _log(f"[config3/2] {base.ni}x{base.nj}x{base.nk}, run lengths {nts}")
    st = model.Setup(base)
    bat = model.Batch(st, [(None if s == 0 else [fields[s]], nts[s]) for s in range(4)])
    bat.run(max(nts))
    s = int(np.argmax(nts))
    sc = copy.copy(base)
    sc.faults = [copy.copy(base.faults[0])]
    sc.faults[0].fields = fields[s]
    orc = oracle.Oracle.from_scenario(sc)
    orc.run(nts[s])
    _check_vs_oracle(bat.read_wavefield(s), bat.read_fault(s), orc, sc, f"config3/2 scenario {s}")
    orc.close()
    bat.close()
    st.close()


def test_one_call_drop_in(built_libs):
    """cgfd_drv_rk_curv_col_allstep (drv_rk_curv_col.h:18-42) on a real
    descriptor: the same final state as create / run / read, bit for bit,
    and within the parity tolerance of the oracle."""
    from cgfd import configs, lib as L, model
    import oracle
    sc = configs.small(imethod=2, curvilinear=True)
    sc.nt_total = 31
    st = model.Setup(sc)
    lib = L.load_library()
    nz, ny, nx = sc.shape_ext
    w1 = np.zeros((9, nz, ny, nx), np.float32)
    f1 = np.zeros((1, 11, nz, ny), np.float32)
    p1 = np.zeros((15, ny, nx), np.float32)
    d = L.Desc.from_buffer_copy(st.desc)
    L.check(lib, lib.cgfd_drv_rk_curv_col_allstep(C.byref(d), L.ptr(w1), L.ptr(f1), L.ptr(p1)))
    core = model.Core(st)
    core.run(sc.nt_total)
    assert np.array_equal(w1, core.read_wavefield())
    assert np.array_equal(f1[0], core.read_fault(0))
    assert np.array_equal(p1, core.read_pg())
    core.close()
    orc = oracle.Oracle.from_scenario(sc)
    orc.run(sc.nt_total)
    _check_vs_oracle(w1, f1[0], orc, sc, "allstep")
    orc.close()
    st.close()


@pytest.mark.parametrize("batch", [False, True])
def test_autotune_keeps_state_and_is_recorded(built_libs, batch):
    """Autotuning between steps leaves the state bit-identical to a run that
    sets the chosen schedule by hand; a new core of the same shape starts
    with the chosen schedule; the chosen schedule passes parity."""
    from cgfd import configs, model
    import oracle
    sc = configs.small(ni=70, nj=52, nk=40, imethod=2, curvilinear=True)
    st = model.Setup(sc)

    def make():
        if batch:
            from test_gpu_batch import _scenario_fields
            return model.Batch(st, [(None, 40), (_scenario_fields(sc, 5), 40)])
        return model.Core(st)

    a, b = make(), make()
    a.run(10)
    b.run(10)
    v, z, ms = a.autotune(2)
    _log(f"[autotune batch={batch}] variant {v}, zchunk {z}, {ms:.3f} ms per launch")
    b.set_variant(v)
    b.set_zchunk(z)
    a.run(15)
    b.run(15)
    ra = a.read_wavefield(0) if batch else a.read_wavefield()
    rb = b.read_wavefield(0) if batch else b.read_wavefield()
    assert np.array_equal(ra, rb)
    c = make()
    assert c.variant_info() == (v, z)
    c.close()
    if not batch:
        orc = oracle.Oracle.from_scenario(sc)
        orc.run(25)
        _check_vs_oracle(ra, a.read_fault(0), orc, sc, f"autotuned variant {v}")
        orc.close()
    a.close()
    b.close()
    st.close()
