"""Batched ensembles on the GPU: scenarios with their own friction and
stress fields and their own run lengths, on one shared grid, against the
same scenarios run one by one through the plain core (and the oracle for
one of them)."""
import numpy as np
import pytest

from helpers import fault_phys, interior, rel_l2

pytestmark = pytest.mark.gpu


def _scenario_fields(sc, seed):
    """Perturbed copies of the configuration's fault inputs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    f = {k: v.copy() for k, v in sc.faults[0].fields.items()}
    f["Ty"] = f["Ty"] * np.float32(1.0 + 0.02 * rng.uniform(-1, 1))
    f["Dc"] = f["Dc"] * np.float32(rng.uniform(0.8, 1.2))
    f["mu_d"] = f["mu_d"] + np.float32(rng.uniform(-0.02, 0.02))
    return [f]


def _plain_with_fields(sc, fields):
    from cgfd import model
    import copy
    sc2 = copy.copy(sc)
    sc2.faults = [copy.copy(sc.faults[0])]
    sc2.faults[0].fields = fields[0]
    return model.Setup(sc2)


CASES = [(v, "default") for v in (1, 20, 23, 24, 32, 33)] + [(20, "first_retires"), (1, "first_retires"),
                                                          (24, "first_retires"), (32, "first_retires"),
                                                          (23, "ragged_slab"), (20, "ragged_slab"),
                                                          (24, "ragged_slab"), (32, "ragged_slab"),
                                                          (23, "wide"), (24, "wide"), (33, "wide"), (32, "wide")]


@pytest.mark.parametrize("variant,case", CASES)
def test_batch_matches_plain_runs(built_libs, variant, case):
    """default: 4 scenarios, scenarios 1 and 3 retire first; first_retires:
    scenario 0 retires first and one scenario runs alone at the end;
    ragged_slab: nj = 50, so the last slab of 4-row tiles is partial; wide:
    ni = 110, wide enough for the box split's interior box (columns 32..95)
    next to its bands."""
    from cgfd import configs, model
    import copy
    import oracle
    if case == "ragged_slab":
        sc = configs.small(imethod=2, nj=50)
    elif case == "wide":
        sc = configs.small(imethod=2, ni=110, nj=50)
    else:
        sc = configs.small(imethod=2)
    st = model.Setup(sc)
    nts = [20, 40, 33] if case == "first_retires" else [40, 25, 40, 33]
    fields = [None] + [_scenario_fields(sc, 100 + s) for s in range(1, len(nts))]
    bat = model.Batch(st, list(zip(fields, nts)), variant=variant)
    bat.run(max(nts) + 5)
    bat.sync()
    assert bat.active == 0
    for s, nt in enumerate(nts):
        assert bat.steps(s) == nt
        st_s = st if fields[s] is None else _plain_with_fields(sc, fields[s])
        core = model.Core(st_s, variant=variant)
        core.run(nt)
        wg, wp = bat.read_wavefield(s), core.read_wavefield()
        fg, fp = bat.read_fault(s), core.read_fault(0)
        r = rel_l2(interior(wg), interior(wp))
        assert r <= 1e-6, f"scenario {s}: batched vs plain wavefield rel L2 {r:.3e}"
        assert rel_l2(fault_phys(fg)[6], fault_phys(fp)[6]) <= 1e-6
        if s == 1:
            sc_s = copy.copy(sc)
            sc_s.faults = [copy.copy(sc.faults[0])]
            sc_s.faults[0].fields = fields[s][0]
            orc = oracle.Oracle.from_scenario(sc_s)
            orc.run(nt)
            assert rel_l2(interior(wg), interior(orc.read_wavefield())) <= 1e-3
            assert rel_l2(fault_phys(fg)[6], fault_phys(orc.read_fault())[6]) <= 1e-3
        core.close()
    bat.close()


def test_batch_graph_replay_matches_eager(built_libs):
    from cgfd import configs, model
    sc = configs.small(imethod=2)
    st = model.Setup(sc)
    nts = [37, 21, 37]
    fields = [None] + [_scenario_fields(sc, 200 + s) for s in range(1, len(nts))]
    bg = model.Batch(st, list(zip(fields, nts)))
    be = model.Batch(st, list(zip(fields, nts)))
    be.set_graph(False)
    for n in (5, 16, 3, 16):
        bg.run(n)
        be.run(n)
        for s in range(len(nts)):
            assert np.array_equal(bg.read_wavefield(s), be.read_wavefield(s))
            assert np.array_equal(bg.read_fault(s), be.read_fault(s))


def test_batch_fault_receivers_match_plain_runs(built_libs):
    """Fault receivers on a batch whose scenarios retire at different steps
    (cgfd_core_set_fault_receivers on the batch core, read per scenario with
    cgfd_batch_read_fault_receivers): each scenario holds the records of its
    own steps only, equal to its plain run's receivers (rel L2 1e-6), and
    the scenarios' records differ (their initial tractions do)."""
    from cgfd import configs, model
    sc = configs.small(imethod=2)
    st = model.Setup(sc)
    nts = [40, 25, 40, 33]
    fields = [None] + [_scenario_fields(sc, 100 + s) for s in range(1, len(nts))]
    j1, j2, k1, k2 = sc.faults[0].grid
    fjk = np.array([[0, (j1 + j2) // 2 + 3, (k1 + k2) // 2 + 3], [0, j1 + 5, k2 + 2], [0, j2 + 2, k1 + 5]], np.int32)
    bat = model.Batch(st, list(zip(fields, nts)))
    bat.set_fault_receivers(fjk, 48)
    bat.run(max(nts) + 5)
    bat.sync()
    recs = []
    for s, nt in enumerate(nts):
        rb = bat.read_fault_receivers(s)
        assert rb.shape == (3, nt, 9)
        st_s = st if fields[s] is None else _plain_with_fields(sc, fields[s])
        core = model.Core(st_s)
        core.set_fault_receivers(fjk, 48)
        core.run(nt)
        rp = core.read_fault_receivers()
        core.close()
        assert rp.shape == rb.shape and np.abs(rp).max() > 0
        r = rel_l2(rb, rp)
        assert r <= 1e-6, f"scenario {s}: batched vs plain fault receivers rel L2 {r:.3e}"
        recs.append(rb)
    assert np.abs(recs[0][:, :nts[1]] - recs[1]).max() > 0
    bat.close()
    st.close()
