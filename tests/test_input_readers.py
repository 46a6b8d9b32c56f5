"""The grid, metric and media imports (cgfd_io_read_coord / _metric / _media)
and the setup's imported-metric path, on the CPU.

Files are written by scipy's independent NetCDF classic writer in the
layouts of the reference's imports: coord_<frame>.nc x, y, z over (k, j, i)
(gd_curv_coord_import, gd_t.cu:855-923, reads the leading ni x nj x nk block
from start 0), metric_<frame>.nc jac, xi_x .. zeta_z and media_<frame>.nc
lambda, mu, rho (gd_curv_metric_import, gd_t.cu:1015-1068; md_import,
md_t.cu:229-286: whole variables of exactly nk x nj x ni). A setup built from
the imported files equals the setup built from the same arrays in memory,
and a setup given the metric it computed itself is bit-identical to it.
"""
import ctypes as C

import numpy as np
import pytest

METRIC_NAMES = ("jac", "xi_x", "xi_y", "xi_z", "eta_x", "eta_y", "eta_z", "zeta_x", "zeta_y", "zeta_z")


def _write(path, dims, vars_):
    from scipy.io import netcdf_file
    with netcdf_file(str(path), "w", version=1) as f:
        for n, v in dims:
            f.createDimension(n, v)
        for name, (typ, arr) in vars_.items():
            f.createVariable(name, typ, tuple(n for n, _ in dims))[:] = arr


def _read(fn, path, ni, nj, nk, ncmp):
    from cgfd import lib as L
    lib = L.load_library()
    out = np.zeros((ncmp, nk, nj, ni), np.float32)
    rc = getattr(lib, fn)(str(path).encode(), ni, nj, nk, out.ctypes.data_as(C.c_void_p))
    return rc, out, lib


def test_coord_reader_reads_the_leading_block(built_libs, tmp_path):
    ni, nj, nk = 9, 7, 5
    rng = np.random.default_rng(11)
    big = rng.standard_normal((3, nk + 2, nj + 1, ni + 3)) * 1e3
    path = tmp_path / "coord_px0_py0_pz0.nc"
    _write(path, [("k", nk + 2), ("j", nj + 1), ("i", ni + 3)],
           {"x": ("f", big[0].astype(np.float32)), "y": ("d", big[1]), "z": ("f", big[2].astype(np.float32))})
    rc, out, lib = _read("cgfd_io_read_coord", path, ni, nj, nk, 3)
    assert rc == 0, lib.cgfd_last_error()
    assert np.array_equal(out[0], big[0, :nk, :nj, :ni].astype(np.float32))
    assert np.array_equal(out[1], big[1, :nk, :nj, :ni].astype(np.float32))
    assert np.array_equal(out[2], big[2, :nk, :nj, :ni].astype(np.float32))
    # a variable smaller than the grid is an error, not a short read
    rc, _, lib = _read("cgfd_io_read_coord", path, ni + 4, nj, nk, 3)
    assert rc == -1


@pytest.mark.parametrize("fn, names", [("cgfd_io_read_metric", METRIC_NAMES),
                                       ("cgfd_io_read_media", ("lambda", "mu", "rho"))])
def test_metric_and_media_readers_need_the_exact_shape(built_libs, tmp_path, fn, names):
    ni, nj, nk = 8, 6, 4
    rng = np.random.default_rng(12)
    vals = rng.standard_normal((len(names), nk, nj, ni)).astype(np.float32)
    path = tmp_path / "in.nc"
    _write(path, [("k", nk), ("j", nj), ("i", ni)],
           {n: ("d" if q == 1 else "f", vals[q]) for q, n in enumerate(names)})
    rc, out, lib = _read(fn, path, ni, nj, nk, len(names))
    assert rc == 0, lib.cgfd_last_error()
    assert np.array_equal(out, vals)
    rc, _, lib = _read(fn, path, ni - 1, nj, nk, len(names))
    assert rc == -1 and b"is not [4][6][7]" in lib.cgfd_last_error()
    # a missing variable names itself
    path2 = tmp_path / "short.nc"
    _write(path2, [("k", nk), ("j", nj), ("i", ni)], {n: ("f", vals[q]) for q, n in enumerate(names[:-1])})
    rc, _, lib = _read(fn, path2, ni, nj, nk, len(names))
    assert rc == -1 and names[-1].encode() in lib.cgfd_last_error()


def _setup_arrays(st, sc):
    from test_setup_ref import _as, _fault_arrays
    nz, ny, nx = sc.shape_ext
    out = {"metric": st.metric(), "lambda": _as(st.desc.lambda_, nz * ny * nx), "mu": _as(st.desc.mu, nz * ny * nx),
           "slw": _as(st.desc.rho, nz * ny * nx)}
    for fid in range(len(sc.faults)):
        for k, v in _fault_arrays(st.desc, fid, nz * ny, ny).items():
            if isinstance(v, list):
                for q, a in enumerate(v):
                    out[f"{fid}.{k}{q}"] = a
            else:
                out[f"{fid}.{k}"] = v
    return out


def _assert_same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_setup_from_imported_files_equals_setup_in_memory(built_libs, tmp_path):
    """coord, metric and media files written from a curvilinear scenario and
    its own metric, read back and fed to the setup: the same setup, bit for
    bit, with and without the imported metric."""
    import copy
    from cgfd import configs, model
    sc = configs.small(imethod=2, curvilinear=True)
    ni, nj, nk = sc.ni, sc.nj, sc.nk
    st = model.Setup(sc)
    ref = _setup_arrays(st, sc)
    met_phys = st.metric()[:, 3:-3, 3:-3, 3:-3]
    st.close()
    dims = [("k", nk), ("j", nj), ("i", ni)]
    _write(tmp_path / "coord_px0_py0_pz0.nc", dims, {"x": ("f", sc.x), "y": ("f", sc.y), "z": ("f", sc.z)})
    _write(tmp_path / "metric_px0_py0_pz0.nc", dims, {n: ("f", met_phys[q]) for q, n in enumerate(METRIC_NAMES)})
    _write(tmp_path / "media_px0_py0_pz0.nc", dims, {"lambda": ("f", sc.lam), "mu": ("f", sc.mu), "rho": ("f", sc.rho)})
    rc, xyz, lib = _read("cgfd_io_read_coord", tmp_path / "coord_px0_py0_pz0.nc", ni, nj, nk, 3)
    assert rc == 0, lib.cgfd_last_error()
    rc, met, lib = _read("cgfd_io_read_metric", tmp_path / "metric_px0_py0_pz0.nc", ni, nj, nk, 10)
    assert rc == 0, lib.cgfd_last_error()
    rc, media, lib = _read("cgfd_io_read_media", tmp_path / "media_px0_py0_pz0.nc", ni, nj, nk, 3)
    assert rc == 0, lib.cgfd_last_error()
    s2 = copy.copy(sc)
    s2.x, s2.y, s2.z = xyz[0], xyz[1], xyz[2]
    s2.lam, s2.mu, s2.rho = media[0], media[1], media[2]
    for metric in (None, met):
        st2 = model.Setup(s2, metric=metric)
        _assert_same(_setup_arrays(st2, s2), ref)
        st2.close()


def test_imported_metric_is_used(built_libs):
    """An imported metric that differs from the coordinates' one is what the
    setup keeps (ghosts by linear extrapolation), and it reaches the fault
    coefficients: a uniformly scaled Jacobian scales every D matrix."""
    from cgfd import configs, model
    sc = configs.small(imethod=2, curvilinear=False)
    st = model.Setup(sc)
    ref = _setup_arrays(st, sc)
    met = st.metric()[:, 3:-3, 3:-3, 3:-3].copy()
    st.close()
    met[0] *= 2.0
    st2 = model.Setup(sc, metric=met)
    got = _setup_arrays(st2, sc)
    st2.close()
    assert np.array_equal(got["metric"][0, 3:-3, 3:-3, 3:-3], met[0])
    # ghosts: 2 b[edge] - b[mirror] of an exactly doubled field is doubled
    assert np.array_equal(got["metric"][0], 2.0 * ref["metric"][0])
    assert np.array_equal(got["metric"][1:], ref["metric"][1:])
    for q in range(12):
        assert np.array_equal(got[f"0.D{q}"], 2.0 * ref[f"0.D{q}"]), q
    with pytest.raises(ValueError):
        model.Setup(sc, metric=met[:, :-1])
