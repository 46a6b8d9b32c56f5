"""The adjoint over the differentiation configuration's full run length:
2000 steps on the 160 x 240 x 100 grid (100 m, PML 12, free surface, a
12 x 6 km fault from the surface, 32 surface stations plus 8 beside the
fault). The friction tape holds 16 floats per fault node and step, 3.3 GB
for the 246 x 106 fault slice over 2000 steps, so its late slots lie past
2^31 bytes.

* locked fault (mu_s = 1000): the map from the initial wavefield to the
  2000 steps of station traces is linear, so the reverse sweep, which reads
  every slot of the tape, must satisfy <g, w0> = sum r . T(w0);
* rupturing fault: a residual that is zero after step 40 gives, through the
  2000-step tape, the gradients of a 40-step run with the same residual
  (the late slots contribute exact zeros); the two sweeps differ only in
  the order of the adjoint kernels' floating-point atomics.
"""
import numpy as np
import pytest

from helpers import progress

pytestmark = pytest.mark.gpu

NT = 2000
WINDOW = 40
PARAMS = ["mu_s", "mu_d", "Dc", "C0", "Tx", "Ty", "Tz"]


def _scenario(locked):
    from test_gpu_adjoint_diffgrid import _scenario as diffgrid
    sc, frec = diffgrid()
    if locked:
        sc.faults[0].fields["mu_s"][:] = 1000.0
    return sc, frec


def _smooth_field(sc, seed, amp):
    from test_gpu_adjoint_fault import _smooth_field as sf
    return sf(sc, seed, amp)


def test_locked_fault_identity_over_2000_steps(built_libs):
    from cgfd import model
    sc, _ = _scenario(locked=True)
    st = model.Setup(sc)
    core = model.Core(st)
    core.set_stations(sc.stations + 3, NT)
    core.adjoint_forward(NT)
    w0 = _smooth_field(sc, 41, 1e-3)
    core.write_wavefield(w0)
    progress(f"[long tape] locked fault, {NT} forward steps on {sc.ni}x{sc.nj}x{sc.nk}")
    core.run(NT)
    tr = core.read_stations()
    assert np.isfinite(tr).all() and np.abs(tr[:, :, :3]).max() > 0
    rng = np.random.default_rng(42)
    r = rng.standard_normal(tr.shape).astype(np.float32)
    r[:, :, 3:] *= 1e-6
    progress(f"[long tape] locked fault, {NT} reverse steps")
    g, gf = core.adjoint_fault(r)
    lhs = float(np.sum(g.astype(np.float64) * w0.astype(np.float64)))
    rhs = float(np.sum(r.astype(np.float64) * tr.astype(np.float64)))
    rel = abs(lhs - rhs) / abs(rhs)
    progress(f"[long tape] <g, w0> = {lhs:.9e}, sum r.T(w0) = {rhs:.9e}, rel {rel:.2e}")
    assert np.isfinite(g).all() and np.isfinite(gf).all()
    assert rel <= 1e-4
    core.close()
    st.close()


def _window_gradients(st, sc, nt, r_window):
    from cgfd import model
    core = model.Core(st)
    core.set_stations(sc.stations + 3, nt)
    core.adjoint_forward(nt)
    core.run(nt)
    tr = core.read_stations()
    r = np.zeros_like(tr)
    r[:, :WINDOW] = r_window
    g, gf = core.adjoint_fault(r)
    core.close()
    return tr, g.astype(np.float64), gf.astype(np.float64)


def test_rupture_early_window_through_the_2000_step_tape(built_libs):
    sc, _ = _scenario(locked=False)
    from cgfd import model
    st = model.Setup(sc)
    rng = np.random.default_rng(43)
    nsta = len(sc.stations)
    r_w = rng.standard_normal((nsta, WINDOW, 9)).astype(np.float32)
    r_w[:, :, 3:] *= 1e-6
    progress(f"[long tape] rupture, {NT} steps forward and reverse")
    tr_l, g_l, gf_l = _window_gradients(st, sc, NT, r_w)
    progress(f"[long tape] rupture, {WINDOW} steps forward and reverse")
    tr_s, g_s, gf_s = _window_gradients(st, sc, WINDOW, r_w)
    st.close()
    assert np.array_equal(tr_l[:, :WINDOW], tr_s)          # the same forward run
    assert np.abs(tr_l[:, WINDOW:, :3]).max() > np.abs(tr_s[:, :, :3]).max()   # the rupture went on
    assert np.abs(gf_s).max() > 0
    rel_w = np.linalg.norm(g_l - g_s) / np.linalg.norm(g_s)
    progress(f"[long tape] gradients through the {NT}-step tape against a {WINDOW}-step run: "
             f"wavefield rel {rel_w:.2e}")
    assert rel_w <= 1e-5
    for p, name in enumerate(PARAMS):
        ns = np.linalg.norm(gf_s[0, p])
        if ns == 0.0:
            assert not gf_l[0, p].any(), name
            continue
        rel = np.linalg.norm(gf_l[0, p] - gf_s[0, p]) / ns
        progress(f"[long tape] d{name}: rel {rel:.2e}")
        assert rel <= 1e-5, name
