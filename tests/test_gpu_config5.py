"""The batched-adjoint configuration at full size: 8 scenarios of the
config-3 physics on the 320 x 200 x 120 grid with 4 hills (own friction and
stress fields, fault lengths and nucleation patches, run lengths 11-30
steps), run as one batch with every scenario's friction tape, then two of
the scenarios differentiated (cgfd_batch_adj_backward). The gradient of
J_s = 1/2 sum v^2 over the particle velocities recorded at stations beside
each nucleation patch, with respect to Dc, mu_s and the initial shear
traction along a seeded smooth direction, against FP64 central differences
of the FP64 oracle running that scenario alone on its own setup, to 1e-3.
The directions are restricted to the nodes whose friction branch decisions
stay at least 1e-4 from their switch points over the FP64 run (as in
test_gpu_adjoint_fault.py).
"""
import copy

import numpy as np
import pytest

from helpers import progress

pytestmark = pytest.mark.gpu

SCEN = (0, 5)
STEPS = (("Dc", 1e-7), ("mu_s", 1e-6), ("Ty", 70.0))


def _patch_centre(sc, f):
    """(j, k) of the centre of the nucleation patch: the nodes whose shear
    traction the generator set 1 MPa above static strength"""
    excess = f["Ty"].astype(np.float64) - f["mu_s"].astype(np.float64) * np.abs(f["Tx"].astype(np.float64))
    ks, js = np.nonzero(np.abs(excess - 1.0e6) < 1.0e3)
    return int(round(js.mean())), int(round(ks.mean()))


def _stations(sc, fields):
    i0 = sc.faults[0].x_index
    out = []
    for s in SCEN:
        j, k = _patch_centre(sc, fields[s])
        out += [[i0 - 2, j, k], [i0 + 3, j + 2, k - 2], [i0 - 4, j - 3, k + 1]]
    return np.array(out, np.int32) + 3


def _scenario(base, fields, s):
    sc = copy.deepcopy(base)
    sc.faults[0].fields = copy.deepcopy(fields[s])
    return sc


def _j_fp64(sc, name, d, eps, nt, sta):
    """1/2 sum v^2 at the stations over nt recorded steps, FP64 oracle with
    `name` moved by eps d (runs in a worker process)"""
    import oracle
    s2 = copy.deepcopy(sc)
    s2.faults[0].fields[name] = (s2.faults[0].fields[name].astype(np.float64) + eps * d).astype(np.float32)
    orc = oracle.Oracle.from_scenario(s2, double=True)
    J = 0.0
    for n in range(nt):
        x = orc.read_wavefield()
        v = np.array([x[:3, k, j, i] for (i, j, k) in sta])
        J += 0.5 * float(np.sum(v * v))
        if n < nt - 1:
            orc.run(1)
    orc.close()
    return J


def _direction(sc, seed):
    j1, j2, k1, k2 = sc.faults[0].grid
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((k2 - k1 + 1, j2 - j1 + 1))
    for _ in range(3):
        for ax in (0, 1):
            a = (np.roll(a, 1, ax) + 2 * a + np.roll(a, -1, ax)) / 4
    d = np.zeros((sc.nk, sc.nj))
    d[k1:k2 + 1, j1:j2 + 1] = a / np.abs(a).max()
    return d


def test_config5_batched_gradients_vs_fp64(built_libs):
    from cgfd import configs, model
    import fdpool
    from test_gpu_adjoint_fault import PARAMS, _smooth_branch_mask
    base, fields, nts = configs.config5(nscen=8, steps_scale=0.01)
    nt = min(nts)
    sta = _stations(base, fields)
    progress(f"[config5] {base.ni}x{base.nj}x{base.nk}, 8 scenarios, run lengths {nts}, {nt} recorded steps, "
             f"{len(sta)} stations")
    st = model.Setup(base)
    bat = model.Batch(st, [(None if s == 0 else [fields[s]], nts[s]) for s in range(8)])
    bat.set_stations(sta, nt)
    bat.adjoint_forward(nt)
    bat.run(nt)
    grads, masks = {}, {}
    for s in SCEN:
        tr = bat.read_stations(s)
        assert tr.shape[1] == nt and np.abs(tr[:, :, :3]).max() > 1e-4, "the stations should see the rupture"
        r = np.zeros_like(tr)
        r[:, :, :3] = tr[:, :, :3]
        _, gf = bat.adjoint_fault(s, r)
        assert np.isfinite(gf).all()
        grads[s] = gf
        sc = _scenario(base, fields, s)
        masks[s] = _smooth_branch_mask(sc, nt, bat.read_fault(s))
        progress(f"[config5] scenario {s}: gradient done, {int(masks[s].sum())} smooth-branch nodes")
    bat.close()
    st.close()
    jobs, meta = [], []
    for s in SCEN:
        sc = _scenario(base, fields, s)
        for q, (name, eps) in enumerate(STEPS):
            d = _direction(sc, 5100 + 10 * s + q) * masks[s]
            meta.append((s, name, eps, d))
            jobs += [(sc, name, d, eps, nt, sta), (sc, name, d, -eps, nt, sta)]
    progress(f"[config5] {len(jobs)} FP64 oracle runs of {nt} steps ...")
    Js = fdpool.run_all(_j_fp64, jobs, workers=4)
    worst = 0.0
    for n, (s, name, eps, d) in enumerate(meta):
        f0 = fields[s][name].astype(np.float64)
        step = (f0 + eps * d).astype(np.float32).astype(np.float64) - (f0 - eps * d).astype(np.float32).astype(np.float64)
        ad = float(np.sum(grads[s][0, PARAMS.index(name)].astype(np.float64) * step))
        fd = Js[2 * n] - Js[2 * n + 1]
        rel = abs(ad - fd) / max(abs(fd), 1e-300)
        worst = max(worst, rel)
        progress(f"[config5] scenario {s} d{name}: batched adjoint {ad:.8e}, FP64 central difference {fd:.8e}, "
                 f"rel {rel:.2e}")
    assert worst <= 1e-3
