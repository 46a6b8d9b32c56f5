"""Batched adjoint: three rupture scenarios with their own friction and
stress fields in one batch (per-scenario friction tapes in the batched
forward run, cgfd_batch_adj_backward per scenario) against each scenario run
and differentiated alone through a single-scenario core. The batched
forward differs from the plain one only in summation order (about 1e-6), so
the gradients agree along seeded smooth directions to 1e-3, and as fields to
a few percent (nodes whose friction branch the two trajectories take at
different steps)."""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NT = 30
PARAMS = ["mu_s", "mu_d", "Dc", "C0", "Tx", "Ty", "Tz"]


def _smooth(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape)
    for _ in range(2):
        for ax in range(a.ndim):
            a = (np.roll(a, 1, ax) + 2 * a + np.roll(a, -1, ax)) / 4
    return a / np.abs(a).max()


def _fields(sc, s):
    f = {k: v.copy() for k, v in sc.faults[0].fields.items()}
    if s == 0:
        return f
    j1, j2, k1, k2 = sc.faults[0].grid
    m = np.zeros_like(f["Dc"], dtype=np.float64)
    m[k1:k2 + 1, j1:j2 + 1] = _smooth((k2 - k1 + 1, j2 - j1 + 1), 300 + s)
    f["Dc"] = (f["Dc"] * (1.0 + 0.2 * m)).astype(np.float32)
    f["mu_d"] = (f["mu_d"] - 0.02 * s * np.abs(m)).astype(np.float32)
    return f


def test_batched_adjoint_matches_single_scenario(built_libs):
    from cgfd import configs, model
    sc = configs.small(imethod=2)
    st = model.Setup(sc)
    fields = [_fields(sc, s) for s in range(3)]
    bat = model.Batch(st, [(None, NT + 5), ([fields[1]], NT + 5), ([fields[2]], NT + 5)])
    bat.set_stations(sc.stations + 3, NT)
    bat.adjoint_forward(NT)
    bat.run(NT)
    dirs = [_smooth((sc.nk, sc.nj), 400 + q) for q in range(3)]
    for s in range(3):
        s1 = copy.deepcopy(sc)
        s1.faults[0].fields = fields[s]
        st1 = model.Setup(s1)
        core = model.Core(st1)
        core.set_stations(s1.stations + 3, NT)
        core.adjoint_forward(NT)
        core.run(NT)
        tr1 = core.read_stations()
        trb = bat.read_stations(s)
        rel_tr = float(np.linalg.norm(trb - tr1) / np.linalg.norm(tr1))
        r = tr1.copy()
        g1w, g1 = core.adjoint_fault(r)
        gbw, gb = bat.adjoint_fault(s, r)
        worst = 0.0
        for q in (0, 1, 2, 5):  # mu_s, mu_d, Dc, shear traction
            for d in dirs:
                a1 = float(np.sum(g1[0, q].astype(np.float64) * d))
                ab = float(np.sum(gb[0, q].astype(np.float64) * d))
                worst = max(worst, abs(ab - a1) / max(abs(a1), 1e-30))
        fld = max(float(np.linalg.norm(gb[0, q] - g1[0, q]) / np.linalg.norm(g1[0, q])) for q in (0, 1, 2, 5))
        wrel = float(np.linalg.norm(gbw - g1w) / np.linalg.norm(g1w))
        print(f"[batch adjoint] scenario {s}: traces {rel_tr:.2e}, directional {worst:.2e}, fields {fld:.2e}, "
              f"initial-state gradient {wrel:.2e}", flush=True)
        assert rel_tr <= 1e-4
        assert worst <= 1e-3
        assert fld <= 5e-2
        assert wrel <= 1e-3
        core.close()
        st1.close()
    bat.close()
    st.close()
