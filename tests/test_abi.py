"""The C ABI without a GPU: the library loads, exports every function the
public headers declare, and rejects bad arguments with a status code and a
message (never exit()). No call here touches a device."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
HEADERS = sorted((REPO / "include").glob("*.h"))

DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(cgfd_\w+)\s*\(", re.M)


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in DECL.finditer(text):
            line = text[m.start():text.find("\n", m.start())]
            if "typedef" in line or "#define" in line:
                continue
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_boundary():
    names = declared_functions()
    for must in ("cgfd_core_create", "cgfd_core_run", "cgfd_core_destroy", "cgfd_drv_rk_curv_col_allstep",
                 "cgfd_setup_create", "cgfd_last_error"):
        assert must in names


def test_library_exports_every_declared_function(built_libs):
    lib = C.CDLL(str(built_libs[0]))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_version_and_error_strings(built_libs):
    from cgfd import lib as L
    lib = L.load_library()
    assert b"gfx950" in lib.cgfd_version()
    assert isinstance(lib.cgfd_last_error(), bytes)


def test_null_arguments_are_rejected(built_libs):
    from cgfd import lib as L
    lib = L.load_library()
    h = C.c_void_p()
    assert lib.cgfd_core_create(None, C.byref(h)) == -1
    assert lib.cgfd_last_error()
    assert lib.cgfd_core_run(None, 1) == -1
    assert lib.cgfd_core_read_wavefield(None, None) == -1
    assert lib.cgfd_core_set_variant(None, 1) == -1
    assert lib.cgfd_setup_create(None, C.byref(h)) == -1
    assert lib.cgfd_drv_rk_curv_col_allstep(None, None, None, None) == -1


def _params(sc):
    """SetupParams for a scenario plus the arrays it points to."""
    from cgfd import lib as L
    from cgfd.model import FAULT_FIELDS
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, np.float32)
        keep.append(a)
        return L.ptr(a)

    p = L.SetupParams()
    p.ni, p.nj, p.nk = sc.ni, sc.nj, sc.nk
    p.x, p.y, p.z = arr(sc.x), arr(sc.y), arr(sc.z)
    p.lambda_, p.mu, p.rho = arr(sc.lam), arr(sc.mu), arr(sc.rho)
    p.free_top = int(sc.free_top)
    for d in range(3):
        for s in range(2):
            p.pml_layers[d][s] = sc.pml_layers[d][s]
            p.pml_alpha_max[d][s], p.pml_beta_max[d][s], p.pml_velocity[d][s] = 3.14, 2.0, 7000.0
    p.imethod = sc.imethod
    p.number_fault = len(sc.faults)
    for fid, f in enumerate(sc.faults):
        p.fault_x_index[fid] = f.x_index
        for q in range(4):
            p.fault_grid[4 * fid + q] = f.grid[q]
        for q, nm in enumerate(FAULT_FIELDS):
            p.init_stress[fid][q] = arr(f.fields[nm])
    p.united_margin = sc.united_margin
    p.dt, p.nt_total = sc.dt, 4
    return p, keep


@pytest.mark.parametrize("mutate,needle", [
    (lambda p: setattr(p, "number_fault", 6), "number_fault"),
    (lambda p: setattr(p, "imethod", 3), "method"),
    (lambda p: p.pml_layers[0].__setitem__(0, 400), "PML"),
    (lambda p: setattr(p, "ni", 0), ""),
])
def test_setup_rejects_bad_parameters(built_libs, mutate, needle):
    from cgfd import configs, lib as L
    lib = L.load_library()
    sc = configs.small()
    p, keep = _params(sc)
    mutate(p)
    h = C.c_void_p()
    rc = lib.cgfd_setup_create(C.byref(p), C.byref(h))
    assert rc == -1
    msg = lib.cgfd_last_error().decode()
    assert needle in msg
    assert not h.value


def test_setup_descriptor_is_consistent(built_libs):
    """The descriptor the setup hands to the core: sizes, PML ranges inside
    the grid, 1/rho in place of rho (main_curv_col_el_3d.cu:705), fault
    masks in {0,1}."""
    from cgfd import configs, model
    sc = configs.small()
    st = model.Setup(sc)
    d = st.desc
    assert (d.ni, d.nj, d.nk) == (sc.ni, sc.nj, sc.nk)
    nz, ny, nx = sc.shape_ext
    n = nx * ny * nz
    rho_inv = np.ctypeslib.as_array(C.cast(d.rho, C.POINTER(C.c_float)), shape=(n,)).reshape(nz, ny, nx)
    assert np.allclose(rho_inv[3:-3, 3:-3, 3:-3], 1.0 / sc.rho, rtol=1e-6)
    for dim in range(3):
        for side in range(2):
            if d.pml_is_sides[dim][side]:
                assert 3 <= d.pml_ni1[dim][side] <= d.pml_ni2[dim][side] <= sc.ni + 2
                assert 3 <= d.pml_nk1[dim][side] <= d.pml_nk2[dim][side] <= sc.nk + 2
    f = d.fault[0]
    ss = ny * nz
    united = np.ctypeslib.as_array(C.cast(f.united, C.POINTER(C.c_int)), shape=(ss,))
    assert set(np.unique(united)) <= {0, 1}
    assert (united == 0).sum() > 0
    st.close()


def test_fault_frame_is_orthonormal(built_libs):
    """vec_n, vec_s1, vec_s2 of the fault plane form an orthonormal frame at
    every node (fault_info.cu:637-700), on topography too."""
    from cgfd import configs, model
    sc = configs.small(curvilinear=True)
    st = model.Setup(sc)
    nz, ny, nx = sc.shape_ext
    ss = ny * nz
    lib = st.lib
    n = np.ctypeslib.as_array(lib.cgfd_setup_vec_n(st.handle, 0), shape=(ss * 3,)).reshape(ss, 3).astype(np.float64)
    f = st.desc.fault[0]
    s1 = np.ctypeslib.as_array(C.cast(f.vec_s1, C.POINTER(C.c_float)), shape=(ss * 3,)).reshape(ss, 3)
    s2 = np.ctypeslib.as_array(C.cast(f.vec_s2, C.POINTER(C.c_float)), shape=(ss * 3,)).reshape(ss, 3)
    # physical nodes only
    g = np.zeros((nz, ny), bool)
    g[3:-3, 3:-3] = True
    g = g.ravel()
    F = np.stack([n[g], s1[g], s2[g]], axis=1)  # (nodes, 3, 3)
    G = np.einsum("nij,nkj->nik", F, F)
    assert np.allclose(G, np.eye(3)[None], atol=2e-6)
    st.close()


def test_pml_profile_known_values(built_libs):
    """R = 10^-((log10 N - 1)/log10 2 + 4) (bdry_t.cu:291-296): 1e-4 at 10
    layers, 1e-5 at 20, 1e-6 at 40; d_max = -(p_d + 1) Vp ln R / (2 L) with
    the quadratic profile power p_d = 2 (bdry_t.cu:15, 298-302)."""
    from cgfd import lib as L
    lib = L.load_library()
    assert np.isclose(lib.cgfd_pml_cal_R(10.0), 1e-4, rtol=1e-5)
    assert np.isclose(lib.cgfd_pml_cal_R(20.0), 1e-5, rtol=1e-5)
    assert np.isclose(lib.cgfd_pml_cal_R(40.0), 1e-6, rtol=1e-5)
    d = lib.cgfd_pml_cal_dmax(2000.0, 6000.0, 1e-5)
    assert np.isclose(d, -3.0 * 6000.0 * np.log(1e-5) / (2.0 * 2000.0), rtol=1e-5)


def test_cfl_dt_max_cartesian(built_libs):
    """blk_t.cu:71-151: dt = CFL * L / Vp with CFL 1.3 and L the smallest
    distance from a node to the planes through its three axis neighbours; on
    a cube lattice of spacing dh that plane is at dh / sqrt(3)."""
    from cgfd import configs, model
    sc = configs.small()
    st = model.Setup(sc)
    assert np.isclose(st.dt_max(), 1.3 * 100.0 / (np.sqrt(3.0) * 6000.0), rtol=1e-5)
    st.close()


def test_product_path_fails_loudly_without_library(tmp_path):
    from cgfd import lib as L
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        L.load_library(tmp_path / "missing.so")
