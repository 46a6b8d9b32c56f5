"""Checks that pin the CPU oracle against what the mathematics fixes (the
reference ships no golden vectors, and it cannot be built here).

* cubic fields on an affine grid: one RK4 step is exact (interior rhs, every
  metric term, 5-point operators, RK4 coefficients);
* linear slip-weakening friction at and around every kink;
* free-surface traction imaging keeps the surface traction at zero;
* CFS-PML absorbs an outgoing pulse that plain zero ghosts reflect;
* a locked split-node fault transmits a pulse without slip (the trial
  traction enforces velocity continuity) and close to a medium without a
  fault;
* FP32 vs FP64 restatements on a rupture.
"""
import numpy as np
import pytest

import pins
from helpers import fault_phys, interior, rel_l2


def _oracle(sc, double=True, plant=0):
    from cgfd import model
    import oracle
    st = model.Setup(sc)
    return st, oracle.Oracle(st.desc, sc.shape_ext, len(sc.faults), double=double, plant=plant)


@pytest.mark.parametrize("double,tol", [(True, 1e-5), (False, 1e-4)])
def test_cubic_fields_one_step_exact(built_libs, double, tol):
    ni, nj, nk = 36, 34, 32
    A = pins.affine_matrix(100.0)
    dt = 0.004
    sc = pins.affine_scenario(ni, nj, nk, A, dt)
    st, orc = _oracle(sc, double)
    n = pins.unit_vector()
    C = pins.cubic_coefs()
    u, L = pins.u_of_grid(ni, nj, nk, A, n)
    w0 = pins.eval_fields(C, u)
    orc.write_wavefield(w0)
    orc.run(1)
    w1 = orc.read_wavefield()
    ex = pins.eval_fields(pins.exact_step(C, n, L, dt), u)
    m = 3 + 13  # ghosts + points reached by stale ghost values in 4 stages
    sl = (slice(None), slice(m, -m), slice(m, -m), slice(m, -m))
    for c in range(9):
        d_num = w1[sl][c] - w0[sl][c]
        d_ex = ex[sl][c] - w0[sl][c]
        r = rel_l2(d_num, d_ex)
        assert r <= tol, f"component {c}: increment rel L2 {r:.3e}"
    orc.close()
    st.close()


def test_cubic_fields_catch_a_dropped_term(built_libs):
    """The exact-step check is sensitive: removing one term of the momentum
    sum (here the eta-direction contribution of Txy to Vx) in the closed
    form changes the Vx increment by far more than the tolerance."""
    ni, nj, nk = 36, 34, 32
    A = pins.affine_matrix(100.0)
    dt = 0.004
    n = pins.unit_vector()
    C = pins.cubic_coefs()
    u, L = pins.u_of_grid(ni, nj, nk, A, n)
    good = pins.exact_step(C, n, L, dt)
    n_bad = n.copy()
    bad_rhs = pins.rhs_coefs(C, n, L, pins.VP ** 2 * pins.RHO - 2 * pins.VS ** 2 * pins.RHO,
                             pins.VS ** 2 * pins.RHO, pins.RHO)
    # drop n_y dTxy/du from dVx/dt
    bad_rhs[0] -= n_bad[1] * pins._dpoly(C)[8] / L / pins.RHO
    inc_good = pins.eval_fields(good - C, u)[0]
    inc_bad = pins.eval_fields(bad_rhs * dt, u)[0]
    assert rel_l2(inc_bad, inc_good) > 1e-2


def test_metric_of_affine_grid_is_exact(built_libs):
    """gd_curv_metric_cal on r = A (i,j,k): jac = det A, [xi; eta; zeta] = A^-1."""
    ni, nj, nk = 20, 18, 16
    A = pins.affine_matrix(50.0, seed=21)
    sc = pins.affine_scenario(ni, nj, nk, A, 0.004)
    from cgfd import model
    st = model.Setup(sc)
    met = st.metric()
    Ai = np.linalg.inv(A)
    core = (slice(3, -3),) * 3
    assert np.allclose(met[0][core], np.linalg.det(A), rtol=2e-5)
    for r in range(3):
        for c in range(3):
            assert np.allclose(met[1 + 3 * r + c][core], Ai[r, c], rtol=2e-4, atol=2e-7 / 50.0), (r, c)
    st.close()


# ---------------------------------------------------------------- friction
def _fric(Trial, T0, mus=0.6, mud=0.5, slip=0.0, dc=0.4, c0=0.0, jacvec=2.0, double=True):
    import oracle
    vn, s1, s2 = [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]
    v = list(Trial) + [jacvec] + vn + list(T0) + [mus, mud, slip, dc, c0] + s1 + s2
    return oracle.probe_friction(v, double)


def test_friction_locked_below_strength(built_libs):
    # local traction (-100, 30, 40): |Ts| = 50 < tau_c = 0.6*100 = 60 -> locked
    out = _fric([0.0, 0.0, 0.0], [-100.0, 30.0, 40.0])
    assert out[6] == 0
    assert np.allclose(out[:3], 0.0)          # T1 = Trial when nothing changes
    assert np.allclose(out[3:6], [-100.0, 30.0, 40.0])


def test_friction_capped_at_strength(built_libs):
    # |Ts| = 100 > 60: shear capped to 60 along the same direction
    out = _fric([0.0, 0.0, 0.0], [-100.0, 60.0, 80.0])
    assert out[6] == 1
    assert np.allclose(out[3:6], [-100.0, 36.0, 48.0])
    # T1 = (Tau - T0) * jacvec
    assert np.allclose(out[:3], [0.0, (36.0 - 60.0) * 2.0, (48.0 - 80.0) * 2.0])


def test_friction_exactly_at_strength_slips(built_libs):
    # |Ts| == tau_c exactly: the comparison is >=, so the node is flagged and Tau = local traction
    out = _fric([0.0, 0.0, 0.0], [-100.0, 36.0, 48.0])
    assert out[6] == 1
    assert np.allclose(out[3:6], [-100.0, 36.0, 48.0])


def test_friction_weakening_kink(built_libs):
    T0 = [-100.0, 0.0, 200.0]
    half = _fric([0.0] * 3, T0, slip=0.2)[5]   # mu_f = 0.55
    at = _fric([0.0] * 3, T0, slip=0.4)[5]     # mu_f = mu_d
    past = _fric([0.0] * 3, T0, slip=0.9)[5]   # mu_f = mu_d
    assert np.isclose(half, 55.0) and np.isclose(at, 50.0) and np.isclose(past, 50.0)


def test_friction_no_opening_and_small_shear_guard(built_libs):
    # tension: normal traction set to zero; |Ts| = 0.05 < 0.1 -> Tau = 0 * n
    out = _fric([0.0] * 3, [10.0, 0.03, 0.04])
    assert out[3] == 0.0 and np.allclose(out[4:6], 0.0)
    assert np.allclose(out[:3], [-10.0 * 2, -0.03 * 2, -0.04 * 2])
    # tension with cohesion 5 and |Ts| = 50: shear capped at C0
    out = _fric([0.0] * 3, [10.0, 30.0, 40.0], c0=5.0)
    assert out[6] == 1 and np.allclose(out[4:6], [3.0, 4.0])


def test_friction_trial_scaling(built_libs):
    # Trial / jacvec enters the local traction: Trial = (0, 20, 0), jacvec 2 -> +10 in Ty
    out = _fric([0.0, 20.0, 0.0], [-100.0, 20.0, 0.0])
    assert out[6] == 0 and np.allclose(out[4], 30.0)


def test_friction_fp32_matches_fp64(built_libs):
    rng = np.random.Generator(np.random.PCG64(9))
    for _ in range(200):
        T0 = [-rng.uniform(1, 200), rng.uniform(-150, 150), rng.uniform(-150, 150)]
        tr = list(rng.uniform(-50, 50, 3))
        slip = rng.uniform(0, 1)
        a = _fric(tr, T0, slip=slip, double=True)
        b = _fric(tr, T0, slip=slip, double=False)
        assert a[6] == b[6]
        assert np.allclose(a[:6], b[:6], rtol=1e-5, atol=1e-4)


# ---------------------------------------------------------------- free surface
@pytest.mark.parametrize("curv", [False, True])
def test_free_surface_stays_traction_free(built_libs, curv):
    """After a rupture reaches the surface, sigma . n on the top layer stays
    at round-off level away from the fault plane (traction imaging, vlow and
    the free-surface rows of the fault kernels)."""
    from cgfd import configs
    sc = configs.small(imethod=2, curvilinear=curv)
    st, orc = _oracle(sc, True)
    orc.run(60)
    w = orc.read_wavefield()
    met = st.metric()
    k = sc.nk - 1 + 3
    i0 = sc.faults[0].x_index + 3
    zx, zy, zz = met[7, k], met[8, k], met[9, k]
    nrm = np.sqrt(zx * zx + zy * zy + zz * zz)
    n = (zx / nrm, zy / nrm, zz / nrm)
    S = w[3:, k]
    txx, tyy, tzz, tyz, txz, txy = S
    t = [txx * n[0] + txy * n[1] + txz * n[2], txy * n[0] + tyy * n[1] + tyz * n[2],
         txz * n[0] + tyz * n[1] + tzz * n[2]]
    mask = np.ones(txx.shape, bool)
    mask[:, i0 - 0:i0 + 1] = False
    mask[:3, :] = mask[-3:, :] = False
    mask[:, :3] = mask[:, -3:] = False
    tmax = max(np.abs(ti[mask]).max() for ti in t)
    smax = np.abs(w[3:, k - 3]).max()
    assert smax > 1e4, "the case must load the near-surface layers"
    assert tmax <= 1e-6 * smax, f"surface traction {tmax:.3e} vs stress {smax:.3e}"
    orc.close()
    st.close()


# ---------------------------------------------------------------- CFS-PML
def _pulse_energy(pml: int, steps: int):
    ni = nj = nk = 40
    A = np.diag([100.0, 100.0, 100.0])
    sc = pins.affine_scenario(ni, nj, nk, A, 0.008, pml=pml)
    st, orc = _oracle(sc, False)
    nz, ny, nx = sc.shape_ext
    kk, jj, ii = np.meshgrid(np.arange(nz) - 3.0, np.arange(ny) - 3.0, np.arange(nx) - 3.0, indexing="ij")
    r2 = (ii - 19.5) ** 2 + (jj - 19.5) ** 2 + (kk - 19.5) ** 2
    w = np.zeros((9, nz, ny, nx))
    w[0] = np.exp(-r2 / (2 * 2.0 ** 2))
    w[2] = 0.5 * np.exp(-r2 / (2 * 2.0 ** 2))
    orc.write_wavefield(w)

    def energy():
        v = orc.read_wavefield()
        core = (slice(None),) + (slice(3 + 10, -3 - 10),) * 3
        return float((v[0:3][core] ** 2).sum())

    e0 = energy()
    orc.run(steps)
    e1 = energy()
    orc.close()
    st.close()
    return e0, e1


def test_cfspml_absorbs_outgoing_pulse(built_libs):
    e0, e_pml = _pulse_energy(10, 110)
    _, e_none = _pulse_energy(0, 110)
    assert np.isfinite(e_pml)
    assert e_pml < 0.01 * e0, f"PML leaves {e_pml / e0:.3e} of the kinetic energy"
    assert e_none > 20 * e_pml, "zero ghosts must reflect far more than the PML"


# ---------------------------------------------------------------- locked fault
def _locked_fault_run(imethod, with_fault, sigma, steps=60, plant=0):
    """FP64 oracle: a Gaussian velocity pulse 8 nodes in front of a fault
    whose static friction is 1000 everywhere (never reaches strength)."""
    from cgfd import configs
    sc = configs.small(imethod=imethod)
    sc.faults[0].fields["mu_s"][:] = 1000.0
    i0 = sc.faults[0].x_index + 3
    if not with_fault:
        sc.faults = []
    st, orc = _oracle(sc, True, plant=plant)
    nz, ny, nx = sc.shape_ext
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    g = np.exp(-(((ii - (i0 - 8)) / sigma) ** 2 + ((jj - ny // 2) / sigma) ** 2 + ((kk - nz // 2) / sigma) ** 2))
    w0 = np.zeros((9, nz, ny, nx))
    w0[0], w0[1], w0[2] = g, 0.5 * g, -0.3 * g
    w0[:, :3] = w0[:, -3:] = 0
    w0[:, :, :3] = w0[:, :, -3:] = 0
    w0[:, :, :, :3] = w0[:, :, :, -3:] = 0
    orc.write_wavefield(w0)
    vmax = vsmax = 0.0
    for _ in range(steps // 5):
        orc.run(5)
        if with_fault:
            w = orc.read_wavefield()
            vmax = max(vmax, float(np.abs(w[:3, :, :, i0 - 3:i0 + 4]).max()))
            vsmax = max(vsmax, float(np.abs(orc.read_fault(0)[3]).max()))
    w = orc.read_wavefield()
    orc.close()
    st.close()
    return w, vmax, vsmax


@pytest.mark.parametrize("imethod", [1, 2])
def test_locked_fault_transmits_without_slip(built_libs, imethod):
    """The trial traction (trial_slipweakening.cu:91-275) is the traction that
    keeps the two sides' velocities equal after the stage; below strength it
    is applied as is, so the split-node update (fault_gpu.cu:326-425,
    fault_wav_t.cu:318-366) must keep the slip rate at rounding level while
    the pulse crosses the plane. A wrong sign, mass or image coefficient in
    the trial assembly or in the split-node velocity rhs breaks the
    continuity and opens O(1) slip."""
    w, vmax, vsmax = _locked_fault_run(imethod, True, 2.5)
    assert vmax > 0.05, "the pulse must reach the fault"
    assert vsmax <= 1e-12 * vmax, f"slip rate {vsmax:.3e} on a locked fault (particle velocity {vmax:.3e})"


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_locked_fault_catches_planted_mistakes(built_libs, plant):
    """The continuity check fails for a trial traction without its factor 2
    (1), with the wrong sign of the side term (2), and for a split-node
    velocity row without its a_0 term (3): each opens O(0.1-1) slip."""
    _, vmax, vsmax = _locked_fault_run(2, True, 2.5, steps=30, plant=plant)
    assert vsmax > 1e-3 * vmax, f"planted mistake {plant} not caught: {vsmax:.3e} vs {vmax:.3e}"


def test_locked_fault_close_to_no_fault(built_libs):
    """The split-node columns use lower-order operators next to the plane,
    so a locked fault differs from the unfaulted medium by truncation error
    only, shrinking as the pulse widens: 14% relative L2 in velocity at a
    width of 2.5 nodes, 8.4% at 4, 6.7% at 6 on this grid, against O(1) for
    a sign or stencil error in the column kernels (fault_gpu.cu:191-1155)."""
    wf, _, _ = _locked_fault_run(2, True, 4.0)
    wn, _, _ = _locked_fault_run(2, False, 4.0)
    r = rel_l2(interior(wf)[:3], interior(wn)[:3])
    assert r < 0.12, f"locked fault vs no fault: velocity rel L2 {r:.3f}"


# ---------------------------------------------------------------- FP32 vs FP64
def test_fp32_restatement_tracks_fp64(built_libs):
    from cgfd import configs
    sc = configs.small(imethod=2)
    st, o64 = _oracle(sc, True)
    import oracle
    o32 = oracle.Oracle(st.desc, sc.shape_ext, 1, double=False)
    o64.run(60)
    o32.run(60)
    assert rel_l2(interior(o32.read_wavefield()), interior(o64.read_wavefield())) < 1e-5
    f32, f64 = fault_phys(o32.read_fault()), fault_phys(o64.read_fault())
    assert (f64[10] > 0).sum() > 50
    assert rel_l2(f32[6], f64[6]) < 1e-5
    o32.close()
    o64.close()
    st.close()
