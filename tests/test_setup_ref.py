"""The host setup (metric, media, CFS-PML profiles, fault coefficients and
masks) against an independent FP64 restatement (oracle/setup_ref.py) and
against closed forms; the oracle run on either setup.

Closed forms (fault_info.cu:97-850, gd_t.cu:209-304):
* Cartesian grid of spacing h, isotropic medium: jac = h^3, grad xi =
  (1/h, 0, 0) etc.; D11 = h diag(l+2m, m, m), D12 = h [[0,l,0],[m,0,0],
  [0,0,0]], D21 = h [[0,m,0],[l,0,0],[0,0,0]], matT1toVx = inv(D11),
  matVx2Vz = -[[0,0,1],[0,0,0],[l/(l+2m),0,0]] (the traction-free condition
  dVx/dz = -dVz/dx, dVz/dz = -l/(l+2m) dVx/dx), n = x, s1 = y, s2 = z;
* any affine grid: D11 matT1toVx = I, D11 matVytoVx = D12, D33 matVx2Vz =
  -D31, inv(D11) D11 = I for method 1, s1 parallel to dr/deta, (n, s1, s2)
  orthonormal, Tn0 = T0 . n.
"""
import numpy as np
import pytest

import pins

G = 3


def _setups(sc):
    from cgfd import model
    from oracle import setup_ref
    return model.Setup(sc), setup_ref.SetupRef(sc)


def _as(ptr, n, dtype=np.float32):
    import ctypes as C
    ct = C.c_float if dtype == np.float32 else C.c_int
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()


def _fault_arrays(desc, fid, ss, ny):
    fd = desc.fault[fid]
    out = {
        "D": [_as(fd.D[q], 9 * ss) for q in range(12)],
        "P2M": [_as(fd.matPlus2Min[q], 9 * ss) for q in range(5)],
        "M2P": [_as(fd.matMin2Plus[q], 9 * ss) for q in range(5)],
        "M2": [_as(fd.matM2[q], 9 * ss) for q in range(6)],
        "F": [_as(fd.matF[q], 9 * ny) for q in range(14)],
        "vec_s1": _as(fd.vec_s1, 3 * ss), "vec_s2": _as(fd.vec_s2, 3 * ss), "rho_f": _as(fd.rho_f, 2 * ss),
        "Tn0": _as(fd.Tn0, ss), "Ts10": _as(fd.Ts10, ss), "Ts20": _as(fd.Ts20, ss),
        "united": _as(fd.united, ss, np.int32), "faultgrid": _as(fd.faultgrid, ss, np.int32),
        "rupy": _as(fd.rup_index_y, ss, np.int32), "rupz": _as(fd.rup_index_z, ss, np.int32),
    }
    return out


def _close(a, b, rtol, what):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    err = np.abs(a - b).max() / scale
    assert err <= rtol, f"{what}: max |diff| / max |ref| = {err:.3e}"


def _compare(sc, rtol=2e-5):
    st, ref = _setups(sc)
    nz, ny, nx = sc.shape_ext
    n = nz * ny * nx
    met = _as(st.desc.metric, 10 * n).reshape(10, nz, ny, nx)
    for c in range(10):
        _close(met[c], ref.met[c], rtol, f"metric {c}")
    _close(_as(st.desc.lambda_, n).reshape(nz, ny, nx), ref.lam, 1e-7, "lambda")
    _close(_as(st.desc.mu, n).reshape(nz, ny, nx), ref.mu, 1e-7, "mu")
    _close(_as(st.desc.rho, n).reshape(nz, ny, nx), ref.slw, 1e-7, "1/rho")
    for (dim, side), f in ref.pml.items():
        assert st.desc.pml_is_sides[dim][side] == int(f["on"])
        r = [st.desc.pml_ni1[dim][side], st.desc.pml_ni2[dim][side], st.desc.pml_nj1[dim][side],
             st.desc.pml_nj2[dim][side], st.desc.pml_nk1[dim][side], st.desc.pml_nk2[dim][side]]
        assert r == f["range"], (dim, side, r, f["range"])
        if f["on"]:
            m = len(f["A"])
            _close(_as(st.desc.pml_A[dim][side], m), f["A"], 1e-5, "pml A")
            _close(_as(st.desc.pml_B[dim][side], m), f["B"], 1e-5, "pml B")
            _close(_as(st.desc.pml_D[dim][side], m), f["D"], 1e-5, "pml D")
    ss = nz * ny
    phys = np.zeros((nz, ny), bool)
    phys[G:-G, G:-G] = True
    for fid, co in enumerate(ref.faults):
        a = _fault_arrays(st.desc, fid, ss, ny)
        for key in ("united", "faultgrid", "rupy", "rupz"):
            assert np.array_equal(a[key].reshape(nz, ny)[phys], co[key][phys]), key
        for key in ("D", "P2M", "M2P", "M2"):
            for q, m in enumerate(co[key]):
                _close(a[key][q].reshape(nz, ny, 9)[phys], m.reshape(nz, ny, 9)[phys], rtol * 10, f"{key}[{q}]")
        # the free-surface row: entries 4..13 go through inv(matVx1_free),
        # whose condition number reaches ~3e4 on topography, so FP32 (the
        # reference's arithmetic) differs from FP64 by ~eps32 * cond there
        cond = np.linalg.cond(co["F"][10][G:-G])
        for q, m in enumerate(co["F"]):
            pa, pb = a["F"][q].reshape(ny, 9)[G:-G].astype(np.float64), m.reshape(ny, 9)[G:-G]
            err = np.abs(pa - pb).max(axis=1) / np.maximum(np.abs(pb).max(axis=1), 1e-30)
            tol = rtol * 50 + (2e-7 * cond if q >= 4 else 0.0)
            assert np.all(err <= tol), f"F[{q}]: worst {err.max():.3e} (tolerance {tol[np.argmax(err)]:.3e})"
        for key in ("vec_s1", "vec_s2"):
            _close(a[key].reshape(nz, ny, 3)[phys], co[key][phys], 1e-6, key)
        _close(a["rho_f"].reshape(2, nz, ny)[:, phys], co["rho_f"][:, phys], 1e-7, "rho_f")
        for key in ("Tn0", "Ts10", "Ts20"):
            _close(a[key].reshape(nz, ny)[phys], co[key][phys], 1e-6, key)
    st.close()


@pytest.mark.parametrize("case", ["small_curv_two_faults", "config1", "buried_dip"])
def test_product_setup_matches_independent_restatement(built_libs, case):
    from cgfd import configs
    if case == "small_curv_two_faults":
        sc = configs.small(curvilinear=True, nfault=2)
    elif case == "config1":
        sc = configs.config1()
    else:
        sc = configs.stress_drop("dip", 1, ni=24, nj=40, nk=40)
    _compare(sc)


# ---------------------------------------------------------------- closed forms
def _cart_checks(co, lam, mu, h, k_top):
    """Raises AssertionError where a closed form fails (node (k_top-2, mid))."""
    k, j = k_top - 2, co["D"][0].shape[1] // 2
    l2m = lam + 2 * mu
    D11 = np.linalg.inv(co["M2"][0][k, j])
    assert np.allclose(D11, h * np.diag([l2m, mu, mu]), rtol=1e-6, atol=1e-6 * h * mu)
    assert np.allclose(co["M2"][1][k, j] @ np.zeros(3), 0)
    D12 = D11 @ co["M2"][1][k, j]
    assert np.allclose(D12, h * np.array([[0, lam, 0], [mu, 0, 0], [0, 0, 0]]), rtol=1e-6, atol=1e-6 * h * mu)
    D13 = D11 @ co["M2"][2][k, j]
    assert np.allclose(D13, h * np.array([[0, 0, lam], [0, 0, 0], [mu, 0, 0]]), rtol=1e-6, atol=1e-6 * h * mu)
    D21 = co["D"][0][k, j]
    assert np.allclose(D21, h * np.array([[0, mu, 0], [lam, 0, 0], [0, 0, 0]]), rtol=1e-6, atol=1e-6 * h * mu)
    D22 = co["D"][1][k, j]
    assert np.allclose(D22, h * np.diag([mu, l2m, mu]), rtol=1e-6, atol=1e-6 * h * mu)
    Vx2Vz = co["F"][0][j]
    assert np.allclose(Vx2Vz, -np.array([[0, 0, 1.0], [0, 0, 0], [lam / l2m, 0, 0]]), atol=1e-6)
    assert np.allclose(co["vec_n"][k, j], [1, 0, 0]) and np.allclose(co["vec_s1"][k, j], [0, 1, 0])
    assert np.allclose(co["vec_s2"][k, j], [0, 0, 1])


def test_setup_closed_forms_cartesian(built_libs):
    """Both setups on a Cartesian grid against the closed forms."""
    from cgfd import configs
    sc = configs.stress_drop("strike", 2, ni=24, nj=40, nk=30)
    st, ref = _setups(sc)
    lam, mu = float(sc.lam.flat[0]), float(sc.mu.flat[0])
    h = 100.0
    k_top = G + sc.nk - 1
    _cart_checks(ref.faults[0], lam, mu, h, k_top)
    nz, ny, _ = sc.shape_ext
    a = _fault_arrays(st.desc, 0, nz * ny, ny)
    co = {key: [m.reshape(nz, ny, 3, 3) for m in a[key]] for key in ("D", "M2")}
    co["F"] = [m.reshape(ny, 3, 3) for m in a["F"]]
    co["vec_n"] = np.broadcast_to(np.array([1.0, 0, 0]), (nz, ny, 3))   # the product keeps vec_n internal
    co["vec_s1"] = a["vec_s1"].reshape(nz, ny, 3)
    co["vec_s2"] = a["vec_s2"].reshape(nz, ny, 3)
    _cart_checks(co, lam, mu, h, k_top)
    # metric of a Cartesian grid
    met = ref.met[:, G:-G, G:-G, G:-G]
    assert np.allclose(met[0], h ** 3, rtol=1e-7)
    assert np.allclose(met[1], 1 / h) and np.allclose(met[5], 1 / h) and np.allclose(met[9], 1 / h)
    for c in (2, 3, 4, 6, 7, 8):
        assert np.abs(met[c]).max() < 1e-12
    st.close()


def test_setup_closed_forms_catch_a_planted_mistake(built_libs):
    """lambda and mu exchanged in the off-diagonal traction terms."""
    from cgfd import configs
    from oracle import setup_ref
    sc = configs.stress_drop("strike", 2, ni=24, nj=40, nk=30)
    bad = setup_ref.SetupRef(sc, plant=1)
    with pytest.raises(AssertionError):
        _cart_checks(bad.faults[0], float(sc.lam.flat[0]), float(sc.mu.flat[0]), 100.0, G + sc.nk - 1)


def test_setup_identities_affine(built_libs):
    """Identities that hold on any grid, checked on a sheared affine grid for
    the independent restatement and the product."""
    import fault_pins as fp
    A = pins.affine_matrix(100.0, seed=4)
    sc = fp.strain_rate_scenario(2, A, ni=24, nj=24, nk=20, free_top=True)
    sc.faults[0].fields["Ty"][:] = 3.0e6
    sc.faults[0].fields["Tz"][:] = -2.0e6
    st, ref = _setups(sc)
    co = ref.faults[0]
    k, j = G + sc.nk - 1, G + sc.nj // 2
    I = np.eye(3)
    D11 = np.linalg.inv(co["M2"][0][k - 4, j])
    assert np.allclose(D11 @ co["M2"][0][k - 4, j], I, atol=1e-10)
    assert np.allclose(co["P2M"][0][k - 4, j], I, atol=1e-10)
    D33, D31 = co["D"][5][k, j], co["D"][3][k, j]
    assert np.allclose(D33 @ co["F"][0][j], -D31, rtol=1e-9, atol=1e-9 * np.abs(D31).max())
    s1 = A[:, 1] / np.linalg.norm(A[:, 1])
    # the coordinates are FP32, so the FP64 metric carries their ~1e-7 rounding
    assert np.allclose(co["vec_s1"][k - 4, j], s1, atol=3e-6)
    n = np.linalg.inv(A)[0]
    n /= np.linalg.norm(n)
    B = np.stack([co["vec_n"][k - 4, j], co["vec_s1"][k - 4, j], co["vec_s2"][k - 4, j]])
    assert np.allclose(co["vec_n"][k - 4, j], n, atol=3e-6)
    assert np.allclose(B @ B.T, I, atol=1e-12)
    T0 = np.array([-100.0e6, 3.0e6, -2.0e6])
    assert np.isclose(co["Tn0"][k - 4, j], T0 @ co["vec_n"][k - 4, j], rtol=1e-12)
    assert np.isclose(co["Tn0"][k - 4, j], T0 @ n, rtol=1e-5)
    # the product's FP32 arrays satisfy the same identities to FP32 rounding
    nz, ny, _ = sc.shape_ext
    a = _fault_arrays(st.desc, 0, nz * ny, ny)
    M2 = [m.reshape(nz, ny, 3, 3).astype(np.float64) for m in a["M2"]]
    Dp = [m.reshape(nz, ny, 3, 3).astype(np.float64) for m in a["D"]]
    Fp = [m.reshape(ny, 3, 3).astype(np.float64) for m in a["F"]]
    D11p = np.linalg.inv(M2[0][k - 4, j])
    assert np.allclose(D11p @ M2[1][k - 4, j], D11 @ co["M2"][1][k - 4, j], rtol=1e-4,
                       atol=1e-5 * np.abs(D11).max())
    assert np.allclose(Dp[5][k, j] @ Fp[0][j], -Dp[3][k, j], rtol=1e-4, atol=1e-5 * np.abs(D31).max())
    assert np.allclose(a["vec_s1"].reshape(nz, ny, 3)[k - 4, j], s1, atol=1e-6)
    st.close()


def test_oracle_on_either_setup_agrees(built_libs):
    """The FP64 oracle on the product's descriptor and on the independent
    one: a 40-step rupture on the topographic two-fault case agrees to the
    FP32 rounding of the setups."""
    from cgfd import configs
    import oracle
    from helpers import fault_phys, interior, rel_l2
    sc = configs.small(curvilinear=True, nfault=2)
    st, ref = _setups(sc)
    o1 = oracle.Oracle(st.desc, sc.shape_ext, 2, double=True)
    d2 = ref.desc()
    o2 = oracle.Oracle(d2, sc.shape_ext, 2, double=True)
    o1.run(40)
    o2.run(40)
    assert rel_l2(interior(o2.read_wavefield()), interior(o1.read_wavefield())) < 1e-5
    f1, f2 = fault_phys(o1.read_fault(0)), fault_phys(o2.read_fault(0))
    assert (f1[10] > 0).sum() > 20
    assert rel_l2(f2[6], f1[6]) < 1e-5
    o1.close()
    o2.close()
    st.close()
