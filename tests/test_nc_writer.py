"""The NetCDF classic writer of the output files (host/cgfd_nc.cpp), on the
CPU: a small driver built with g++ writes a record file (unlimited time,
record and fixed variables, integer global attributes, records written out
of order, the fixed variable last), and scipy's independent NetCDF reader
must see the same dimensions, variables, attributes and values."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
HOST = REPO / "gpu-cgfd3d-dynamic_amd" / "host"

DRIVER = r"""
#include "cgfd_nc.h"
#include <cstdio>
#include <vector>
int main(int argc, char **argv) {
  cgfd::NcWriter w;
  if (w.create(argv[1])) return 1;
  int dt = w.def_dim("time", 0), dk = w.def_dim("k", 3), dj = w.def_dim("j", 5);
  int vt = w.def_var("time", {dt});
  int va = w.def_var("Tn", {dt, dk, dj});
  int vb = w.def_var("Vs", {dt, dk, dj});
  int vf = w.def_var("Peak_vs", {dk, dj});
  w.att_int("i_index_with_ghosts_in_this_thread", {21});
  w.att_int("coords_of_mpi_topo", {0, 1, 2});
  if (w.enddef()) return 2;
  const int order[4] = {0, 2, 1, 3};
  for (int q = 0; q < 4; q++) {
    int r = order[q];
    float t = 0.5f * r;
    std::vector<float> a(15), b(15);
    for (int i = 0; i < 15; i++) { a[i] = 100.0f * r + i; b[i] = -1.0f * r - 0.25f * i; }
    if (w.put_rec(vt, r, &t) || w.put_rec(va, r, a.data()) || w.put_rec(vb, r, b.data())) return 3;
  }
  std::vector<float> f(15);
  for (int i = 0; i < 15; i++) f[i] = 1e-3f * i * i;
  if (w.put_var(vf, f.data())) return 4;
  return w.close();
}
"""


@pytest.fixture(scope="module")
def nc_file(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("nc")
    (d / "drv.cpp").write_text(DRIVER)
    exe = d / "drv"
    subprocess.run([gxx, "-O1", "-std=c++17", f"-I{REPO / 'include'}", f"-I{HOST}", str(d / "drv.cpp"),
                    str(HOST / "cgfd_nc.cpp"), str(HOST / "cgfd_errors.cpp"), "-o", str(exe)], check=True)
    path = d / "fault_i21_px0_py0_pz0.nc"
    subprocess.run([str(exe), str(path)], check=True)
    return path


def test_nc_writer_read_back_by_scipy(nc_file):
    from scipy.io import netcdf_file
    with netcdf_file(str(nc_file), "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.dimensions["time"] is None and f.dimensions["k"] == 3 and f.dimensions["j"] == 5
        assert list(f.variables) == ["time", "Tn", "Vs", "Peak_vs"]
        assert f.variables["Tn"].dimensions == ("time", "k", "j")
        assert f.variables["Peak_vs"].dimensions == ("k", "j")
        assert np.array_equal(f.variables["time"][:], np.array([0.0, 0.5, 1.0, 1.5], np.float32))
        for r in range(4):
            a = 100.0 * r + np.arange(15, dtype=np.float32)
            b = -1.0 * r - 0.25 * np.arange(15, dtype=np.float32)
            assert np.array_equal(f.variables["Tn"][r], a.reshape(3, 5).astype(np.float32))
            assert np.array_equal(f.variables["Vs"][r], b.reshape(3, 5).astype(np.float32))
        i = np.arange(15, dtype=np.float32)
        pk = ((np.float32(1e-3) * i) * i).reshape(3, 5)  # the driver's float32 arithmetic
        assert np.array_equal(f.variables["Peak_vs"][:], pk)
        assert int(f.i_index_with_ghosts_in_this_thread) == 21
        assert list(f.coords_of_mpi_topo) == [0, 1, 2]


def test_init_stress_reader(built_libs, tmp_path):
    """init_stress_<n>.nc as the reference's conf_fault_stress.m writes it
    (NetCDF classic, dims nj, nk, float variables over (nk, nj), plus the
    coordinates and increments it also stores), here one of them double:
    read back exactly by cgfd_io_read_init_stress."""
    import ctypes as C
    from scipy.io import netcdf_file
    from cgfd import lib as L
    nj, nk = 17, 9
    rng = np.random.default_rng(3)
    vals = {n: rng.standard_normal((nk, nj)).astype(np.float32) for n in ("Tx", "Ty", "Tz", "mu_s", "mu_d", "Dc", "C0")}
    path = tmp_path / "init_stress_1.nc"
    with netcdf_file(str(path), "w", version=1) as f:
        f.createDimension("nj", nj)
        f.createDimension("nk", nk)
        for n in ("x", "y", "z", "Tx", "Ty", "Tz", "dTx", "dTy", "dTz", "mu_s", "mu_d", "Dc", "C0"):
            v = f.createVariable(n, "d" if n == "C0" else "f", ("nk", "nj"))
            v[:] = vals[n] if n in vals else np.zeros((nk, nj), np.float32)
    lib = L.load_library()
    out = np.zeros((7, nk, nj), np.float32)
    rc = lib.cgfd_io_read_init_stress(str(path).encode(), nj, nk, out.ctypes.data_as(C.c_void_p))
    assert rc == 0, lib.cgfd_last_error()
    for q, n in enumerate(("Tx", "Ty", "Tz", "mu_s", "mu_d", "Dc", "C0")):
        assert np.array_equal(out[q], vals[n]), n
    # the reference reads the leading nk x nj block (nc_get_vara_float from
    # start 0): a smaller grid reads a block, a larger one is an error
    small = np.zeros((7, nk - 2, nj - 3), np.float32)
    rc = lib.cgfd_io_read_init_stress(str(path).encode(), nj - 3, nk - 2, small.ctypes.data_as(C.c_void_p))
    assert rc == 0, lib.cgfd_last_error()
    assert np.array_equal(small[1], vals["Ty"][:nk - 2, :nj - 3])
    rc = lib.cgfd_io_read_init_stress(str(path).encode(), nj + 1, nk, out.ctypes.data_as(C.c_void_p))
    assert rc == -1 and b"smaller than the grid" in lib.cgfd_last_error()


def test_fault_coord_reader(built_libs, tmp_path):
    """fault_coord_<n>.nc as the reference's conf_fault_grid.m writes it
    (dims nj, nk and dim; x, y, z over (nk, nj); vectors over dim)."""
    import ctypes as C
    from scipy.io import netcdf_file
    from cgfd import lib as L
    nj, nk = 12, 7
    rng = np.random.default_rng(4)
    xyz = rng.standard_normal((3, nk, nj)).astype(np.float32) * 1e4
    path = tmp_path / "fault_coord_1.nc"
    with netcdf_file(str(path), "w", version=1) as f:
        f.createDimension("nj", nj)
        f.createDimension("nk", nk)
        f.createDimension("dim", 3)
        for q, n in enumerate(("x", "y", "z")):
            f.createVariable(n, "f", ("nk", "nj"))[:] = xyz[q]
        f.createVariable("vec_n", "f", ("dim",))[:] = np.array([1, 0, 0], np.float32)
    lib = L.load_library()
    out = np.zeros((3, nk, nj), np.float32)
    rc = lib.cgfd_io_read_fault_coord(str(path).encode(), nj, nk, out.ctypes.data_as(C.c_void_p))
    assert rc == 0, lib.cgfd_last_error()
    assert np.array_equal(out, xyz)
