"""Checks at the full single-GPU shape, 1200 x 600 x 400 nodes, where the
oracle is too slow to run: properties that do not depend on size.

* cubic fields on an affine grid advance by exactly one RK4 step of the
  closed form (the same pin as the oracle's, now on every tile, XCD slot and
  block of the real launch shape);
* every volume-kernel variant the autotuner chooses from produces the same
  rupture on the real configuration (tiles, slab orders and z chunks cover
  every point exactly once).
"""
import gc

import numpy as np
import pytest

import pins
from helpers import rel_l2

pytestmark = pytest.mark.gpu

NI, NJ, NK = 600, 1200, 400


def _log(msg):
    print(msg, flush=True)


def test_fullsize_cubic_fields_one_step(built_libs):
    from cgfd import model
    A = pins.affine_matrix(50.0)
    dt = 0.01  # one step: stability is irrelevant, a larger increment keeps FP32 storage error small
    sc = pins.affine_scenario(NI, NJ, NK, A, dt)
    _log(f"[fullsize-cubic] setup {NI}x{NJ}x{NK}")
    st = model.Setup(sc)
    del sc.x, sc.y, sc.z, sc.lam, sc.mu, sc.rho
    core = model.Core(st, variant=0)
    n = pins.unit_vector()
    C = pins.cubic_coefs()
    nz, ny, nx = NK + 6, NJ + 6, NI + 6
    w = np.zeros((9, nz, ny, nx), np.float32)
    w_c = np.array([(NI - 1) / 2.0, (NJ - 1) / 2.0, (NK - 1) / 2.0])
    wv = n @ A
    L = 0.5 * (abs(wv[0]) * NI + abs(wv[1]) * NJ + abs(wv[2]) * NK)
    ii = np.arange(-3, NI + 3, dtype=np.float64)[None, :]
    jj = np.arange(-3, NJ + 3, dtype=np.float64)[:, None]
    base = (wv[0] * (ii - w_c[0]) + wv[1] * (jj - w_c[1])) / L
    for k in range(nz):
        u = base + wv[2] * ((k - 3) - w_c[2]) / L
        w[:, k] = pins.eval_fields(C, u, np.float32)
    _log("[fullsize-cubic] write + 1 step")
    core.write_wavefield(w)
    core.run(1)
    w1 = core.read_wavefield()
    core.close()
    st.close()
    ex = pins.exact_step(C, n, L, dt)
    m = 3 + 13
    worst = 0.0
    for k in range(m, nz - m, 37):  # every 37th plane of the interior
        u = base + wv[2] * ((k - 3) - w_c[2]) / L
        e = pins.eval_fields(ex, u)[:, m:-m, m:-m]
        w0k = pins.eval_fields(C, u)[:, m:-m, m:-m]
        for c in range(9):
            r = rel_l2(w1[c, k, m:-m, m:-m].astype(np.float64) - w0k[c], e[c] - w0k[c])
            worst = max(worst, r)
    _log(f"[fullsize-cubic] worst increment rel L2 {worst:.3e}")
    assert worst <= 1e-3


def test_fullsize_volume_variants_agree(built_libs):
    from cgfd import configs, model
    sc = configs.config2(steps=8)
    st = model.Setup(sc)
    del sc.x, sc.y, sc.z, sc.lam, sc.mu, sc.rho
    ref = None
    for v, z in ((1, 4), (20, 4), (23, 32), (33, 32), (24, 8), (20, 16)):
        _log(f"[fullsize-variants] variant {v} zchunk {z}")
        core = model.Core(st, variant=v)
        core.set_zchunk(z)
        core.run(6)
        w = core.read_wavefield()
        f = core.read_fault(0)
        core.close()
        if ref is None:
            ref = (w, f)
            assert np.isfinite(w).all()
            assert np.abs(w[:3]).max() > 0.0
            continue
        r = rel_l2(w, ref[0])
        _log(f"[fullsize-variants] variant {v}: wavefield rel L2 vs variant 1 = {r:.3e}")
        assert r <= 1e-4
        assert rel_l2(f[6], ref[1][6]) <= 1e-4
        del w, f
        gc.collect()
    st.close()
