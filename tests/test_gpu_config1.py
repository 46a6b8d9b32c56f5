"""The reference's CPU-runnable case run to the end: 80 x 120 x 60 nodes at
100 m (strike 120, normal 80, depth 60), homogeneous half-space, flat free
surface, a vertical 8 x 4 km fault from the surface, 1.5 km nucleation patch,
CFS-PML of 12 layers on five faces, dt 0.008 s, 800 steps + 1 (the
reference's nt_total, main_curv_col_el_3d.cu:479), 16 surface stations.

The GPU core and the FP32 oracle run the same seeded inputs through the
whole rupture; compared with the target's tolerances: relative L2 <= 1e-3
for velocity snapshots every 200 steps, the final wavefield, final slip,
peak slip rate and the station traces; rupture time within one dt where
both ruptured; the ruptured-node sets agree on >= 99.5% of the nodes.
"""
import time

import numpy as np
import pytest

from helpers import fault_phys, interior, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-3
NAMES = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"]


def _log(msg):
    print(msg, flush=True)


def test_config1_full_run(built_libs):
    from cgfd import configs, model
    import oracle
    sc = configs.config1()
    nt = sc.nt_total
    st = model.Setup(sc)
    core = model.Core(st)
    ijk = sc.stations + 3
    core.set_stations(ijk, nt)
    orc = oracle.Oracle.from_scenario(sc)
    every = 4  # oracle station samples (the GPU records every step)
    rec_o = []
    t0 = time.perf_counter()
    done = 0
    for target in (200, 400, 600, nt):
        while done < target:
            n = min(every, target - done)
            w = orc.read_wavefield()
            rec_o.append([w[:, k, j, i] for (i, j, k) in ijk])
            orc.run(n)
            done += n
        core.run(target - core.step_index)
        wg = interior(core.read_wavefield())
        wo = interior(orc.read_wavefield())
        assert np.isfinite(wg).all()
        for c in range(3):
            r = rel_l2(wg[c], wo[c])
            assert r <= TOL, f"step {target}: {NAMES[c]} rel L2 {r:.3e}"
        _log(f"[config1] step {target}: velocity rel L2 "
             f"{max(rel_l2(wg[c], wo[c]) for c in range(3)):.2e} ({time.perf_counter() - t0:.0f} s)")
    for c in range(3, 9):
        r = rel_l2(wg[c], wo[c])
        assert r <= TOL, f"final {NAMES[c]} rel L2 {r:.3e}"
    fg = fault_phys(core.read_fault(0))
    fo = fault_phys(orc.read_fault(0))
    r_slip = rel_l2(fg[6], fo[6])
    r_pv = rel_l2(fg[9], fo[9])
    _log(f"[config1] final slip rel L2 {r_slip:.2e}, peak slip rate rel L2 {r_pv:.2e}, "
         f"max slip {fo[6].max():.3f} m")
    assert fo[6].max() > 0.1, "the case must rupture"
    assert r_slip <= TOL and r_pv <= TOL
    rup_o, rup_g = fo[10] > 0, fg[10] > 0
    assert (rup_o == rup_g).mean() >= 0.995
    both = rup_o & rup_g
    dtt = np.abs(fg[10][both] - fo[10][both])
    assert dtt.max() <= sc.dt * 1.0001 + 1e-6, f"rupture time off by {dtt.max():.4g} s"
    rec_g = core.read_stations()[:, ::every][:, :len(rec_o)]
    rec_o = np.transpose(np.array(rec_o), (1, 0, 2))
    assert rec_g.shape == rec_o.shape
    r_st = rel_l2(rec_g[:, :, :3], rec_o[:, :, :3])
    _log(f"[config1] station velocity traces rel L2 {r_st:.2e}")
    assert r_st <= TOL
    core.close()
    orc.close()
    st.close()
