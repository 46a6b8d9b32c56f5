"""Closed-form pins of the oracle's split-node fault path (FP64), each with
planted-mistake controls that show the pin fails when one term, sign or
coefficient is wrong (the mistakes are planted in the oracle, on the CPU;
see oracle_create in oracle/cgfd_oracle.cpp).

* uniform stress drop (trial_slipweakening.cu:91-388 under slip, the
  split-node velocity rows fault_gpu.cu:191-429, fault2wave transform.cu:
  292-376, the accumulators fault_wav_t.cu:197-264): strike-slip reaching
  the free surface and buried dip-slip, both dynamic methods, against the
  1-D solution;
* uniform strain rate (the split-node stress rows fault_gpu.cu:431-1879 with
  the method-1/2 matrices of fault_info.cu:97-850 and method 2's T1 rate,
  and their free-surface rows): exact after one step on Cartesian and
  sheared grids;
* wave2fault then fault2wave is the identity on a node (with the Txz/Tyz
  argument swap of transform.cu:274-276 unless it is corrected), and the
  fault-plane column of the strain-rate case shows the swap;
* the locked-fault continuity check (test_oracle_pins.py) catches the
  planted trial-assembly mistakes.
"""
import numpy as np
import pytest

import fault_pins as fp
import pins

STRESS_DROP_STEPS = 64


def _oracle(sc, double=True, fix_q1=False, plant=0):
    from cgfd import model
    import oracle
    st = model.Setup(sc)
    return st, oracle.Oracle(st.desc, sc.shape_ext, len(sc.faults), double=double, fix_q1=fix_q1, plant=plant)


def stress_drop_scenario(mode, imethod):
    from cgfd import configs
    if mode == "strike":
        return configs.stress_drop("strike", imethod, ni=40, nj=100, nk=60)
    return configs.stress_drop("dip", imethod, ni=40, nj=100, nk=100)


def _stress_drop(mode, imethod, plant=0):
    sc = stress_drop_scenario(mode, imethod)
    _, _, last = fp.stress_drop_window(sc)
    assert last >= STRESS_DROP_STEPS, "the check window must close before edge waves arrive"
    st, orc = _oracle(sc, True, plant=plant)
    orc.run(STRESS_DROP_STEPS)
    r = fp.check_stress_drop(orc.read_fault(0), orc.read_fault_state(0), orc.read_wavefield(), sc, STRESS_DROP_STEPS)
    orc.close()
    st.close()
    return r


@pytest.mark.parametrize("mode", ["strike", "dip"])
@pytest.mark.parametrize("imethod", [1, 2])
def test_uniform_stress_drop_matches_closed_form(built_libs, mode, imethod):
    r = _stress_drop(mode, imethod)
    assert r["ok"], r


@pytest.mark.parametrize("plant", [3, 6])
def test_uniform_stress_drop_catches_planted_mistakes(built_libs, plant):
    """3: the split-node DxT1 without its a_0 term; 6: fault2wave's plane
    velocity from one side instead of the side average."""
    r = _stress_drop("strike", 2, plant=plant)
    assert not r["ok"], r


# ---------------------------------------------------------------- uniform strain rate
def _affine(cart):
    return np.diag([100.0, 100.0, 100.0]) if cart else pins.affine_matrix(100.0, seed=4)


def strain_rate_run(imethod, cart, free_top, fix_q1=True, plant=0):
    A = _affine(cart)
    sc = fp.strain_rate_scenario(imethod, A, free_top=free_top)
    G = fp.strain_rate_gradient(A, free_top)
    st, orc = _oracle(sc, True, fix_q1=fix_q1, plant=plant)
    w0, f0, t1, w1, f1 = fp.strain_rate_state(sc, A, G)
    orc.write_wavefield(w0)
    orc.write_fault_state(f0, 0, t1)
    orc.run(1)
    out = (orc.read_wavefield(), orc.read_fault_state(0), orc.read_fault(0))
    orc.close()
    st.close()
    return sc, out, (w1, f1)


# Cartesian grids have exact FP32 metrics; the sheared grid's metric comes
# from FP32 coordinates (relative error ~1e-6 after differencing).
SR_TOL = {True: 1e-6, False: 2e-4}


@pytest.mark.parametrize("cart", [True, False])
@pytest.mark.parametrize("free_top", [True, False])
@pytest.mark.parametrize("imethod", [1, 2])
def test_uniform_strain_rate_is_exact(built_libs, cart, free_top, imethod):
    sc, (w, f, fo), (w1, f1) = strain_rate_run(imethod, cart, free_top)
    top = 5 if (free_top and imethod == 2) else 0
    e = fp.strain_rate_errors(w, f, w1, f1, sc, top_skip=top)
    assert max(e) <= SR_TOL[cart], e
    assert np.abs(fo[3]).max() == 0.0 or np.abs(fo[3]).max() < 1e-12   # locked: no slip rate


def test_method2_top_rows_use_no_t1_rate(built_libs):
    """Method 2 takes the T1 rate only inside the fault grid, three nodes in
    from its edges (fault_info.cu:1063-1067, fault_gpu.cu:664-669), so on a
    fault reaching the surface the top three rows of split nodes are not
    exact on a uniform strain rate while method 1's are."""
    sc, (w, f, _), (w1, f1) = strain_rate_run(2, True, True)
    i0 = sc.faults[0].x_index + 3
    top = np.abs(f[3:, :, -4] - f1[3:, :, -4])[..., 17:-17].max() / np.abs(f1[3:]).max()
    assert top > 1e-2
    sc, (w, f, _), (w1, f1) = strain_rate_run(1, True, True)
    top = np.abs(f[3:, :, -4] - f1[3:, :, -4])[..., 17:-17].max() / np.abs(f1[3:]).max()
    assert top < 1e-6
    del i0


@pytest.mark.parametrize("imethod,plant", [(1, 5), (2, 4)])
def test_uniform_strain_rate_catches_planted_mistakes(built_libs, imethod, plant):
    """5: method 1 without the inv(D11) D11 DxV term; 4: method 2 without
    the T1 rate."""
    sc, (w, f, _), (w1, f1) = strain_rate_run(imethod, True, False, plant=plant)
    e = fp.strain_rate_errors(w, f, w1, f1, sc)
    assert max(e) > 1e-2, e


# ---------------------------------------------------------------- wave2fault / fault2wave
def test_wave2fault_fault2wave_round_trip(built_libs):
    import oracle
    A = pins.affine_matrix(100.0, seed=8)
    Ai = np.linalg.inv(A)
    jac = np.linalg.det(A)
    rng = np.random.Generator(np.random.PCG64(31))
    s = rng.uniform(-1.0, 1.0, 6) * 1e6           # xx, yy, zz, yz, xz, xy
    s[3], s[4] = 2.0e6, -3.0e6                    # Txz != Tyz
    back, tr = oracle.probe_w2f_f2w(Ai.reshape(-1), jac, s, fix_q1=True)
    assert np.allclose(back, s, rtol=1e-12, atol=1e-6)
    # the traction matrix is J sigma [grad xi, grad eta, grad zeta]
    S = np.array([[s[0], s[5], s[4]], [s[5], s[1], s[3]], [s[4], s[3], s[2]]])
    assert np.allclose(tr, jac * S @ Ai.T, rtol=1e-12)
    swapped, _ = oracle.probe_w2f_f2w(Ai.reshape(-1), jac, s, fix_q1=False)
    exp = s.copy()
    exp[3], exp[4] = s[4], s[3]
    assert np.allclose(swapped, exp, rtol=1e-12, atol=1e-6)


@pytest.mark.parametrize("fix_q1", [False, True])
def test_plane_column_shows_the_txz_tyz_swap(built_libs, fix_q1):
    """On the sheared grid (sigma_xz != sigma_yz) the unwelded fault-plane
    column carries sigma_xz in the Tyz array and sigma_yz in Txz unless the
    swap is corrected; the rest of the column and the field off the plane
    are exact either way (nothing reads the plane column's stress)."""
    sc, (w, f, _), (w1, f1) = strain_rate_run(1, False, False, fix_q1=fix_q1)
    i0 = sc.faults[0].x_index + 3
    c = (slice(3 + 14, 3 + sc.nk - 14), slice(3 + 14, 3 + sc.nj - 14), i0)
    nrm = np.abs(w1[3:]).max()
    exp_tyz, exp_txz = (w1[6][c], w1[7][c]) if fix_q1 else (w1[7][c], w1[6][c])
    assert np.abs(w[6][c] - exp_tyz).max() <= 2e-4 * nrm
    assert np.abs(w[7][c] - exp_txz).max() <= 2e-4 * nrm
    assert np.abs(w1[6][c] - w1[7][c]).max() > 0.1 * nrm
    for cmp in (3, 4, 5, 8):
        assert np.abs(w[cmp][c] - w1[cmp][c]).max() <= 2e-4 * nrm


def test_kink_margins_of_the_friction_branches(built_libs):
    """The branch margins the adjoint tests select smooth nodes with. Under
    the uniform stress drop (mu_s = mu_d = 0.5, sigma_n = 100 MPa, shear 10
    MPa above strength) the first trial evaluation sees the initial traction,
    so the shear-vs-strength margin is at most 10 / 50 = 0.2, and it is the
    same at every node away from the slipping region's edges; the welded
    nodes never reach the trial (1e300); a locked fault (mu = 1000) keeps
    every node far from its switch."""
    import copy
    import oracle
    sc = stress_drop_scenario("strike", 2)
    orc = oracle.Oracle.from_scenario(sc, double=True)
    orc.run(3)
    m = orc.read_kink_margin(0)
    orc.close()
    act = m[0] < 1e300
    ks, js = np.nonzero(act)
    assert act.sum() > 1000 and np.all(m[:, 0:3, :] == 1e300)
    inner = m[0][ks.min() + 6:ks.max() - 5, js.min() + 6:js.max() - 5]
    assert np.all(inner < 1e300)
    assert inner.max() - inner.min() <= 1e-12
    assert 0.15 < inner.min() <= 0.2 + 1e-12
    assert m[1][act].min() > 0.5        # slip stays far below Dc
    locked = copy.deepcopy(sc)
    locked.faults[0].fields["mu_s"][:] = 1000.0
    locked.faults[0].fields["mu_d"][:] = 1000.0
    orc = oracle.Oracle.from_scenario(locked, double=True)
    orc.run(3)
    ml = orc.read_kink_margin(0)[0]
    orc.close()
    assert ml[ml < 1e300].min() > 0.99
