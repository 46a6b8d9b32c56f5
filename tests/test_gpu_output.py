"""Output files in the reference's NetCDF layouts (include/cgfd_io.h) from a
short rupture: fault frames every 5 steps with Peak_vs/Init_t0 at the end,
x/y/z slices, and the PG file, read back with scipy's NetCDF reader and
compared bit for bit with what the core returns at the same steps, and with
the FP32 oracle run in step with it (the reference's io_fault_nc_put,
io_slice_nc_put and PG_slice_output values at those steps, to the parity
tolerance)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ok(core, rc):
    assert rc == 0, core.lib.cgfd_last_error()


def test_fault_slice_pg_files(built_libs, tmp_path):
    from cgfd import configs, model
    from scipy.io import netcdf_file
    import oracle
    from helpers import rel_l2
    sc = configs.small(imethod=2)
    st = model.Setup(sc)
    core = model.Core(st)
    orc = oracle.Oracle.from_scenario(sc)
    o_faults, o_slc = [], []
    lib, h = core.lib, core.handle
    i0 = sc.faults[0].x_index + 3
    d = str(tmp_path).encode()
    fo = C.c_void_p()
    _ok(core, lib.cgfd_io_fault_create(h, 0, d, b"px0_py0_pz0", C.byref(fo)))
    slices = [(0, i0 - 2), (1, 20), (2, sc.nk + 2)]
    so = []
    for ax, idx in slices:
        o = C.c_void_p()
        _ok(core, lib.cgfd_io_slice_create(h, ax, idx, d, None, C.byref(o)))
        so.append(o)
    box = (4, 9, 3, 5, 7, 5, 6, 8, 2)
    bx = (C.c_int * 9)(*box)
    sn = C.c_void_p()
    _ok(core, lib.cgfd_io_snap_create(h, b"snap_001", bx, 7, d, None, C.byref(sn)))
    dt = float(st.desc.dt)
    faults, slc, snaps = [], [], []
    for it in range(20):
        if it % 5 == 0:
            _ok(core, lib.cgfd_io_fault_put(fo, h, it // 5, C.c_float(it * dt)))
            for o in so:
                _ok(core, lib.cgfd_io_slice_put(o, h, it // 5, C.c_float(it * dt)))
            faults.append(core.read_fault(0)[:9, 3:-3, 3:-3].copy())
            slc.append([core.read_slice(ax, idx) for ax, idx in slices])
            _ok(core, lib.cgfd_io_snap_put(sn, h, it // 5, C.c_float(it * dt)))
            snaps.append(core.read_snapshot(box, 7))
            o_faults.append(orc.read_fault(0)[:9, 3:-3, 3:-3].copy())
            ow = orc.read_wavefield()
            o_slc.append([ow[:, :, :, idx] if ax == 0 else (ow[:, :, idx, :] if ax == 1 else ow[:, idx, :, :])
                          for ax, idx in slices])
        core.run(1)
        orc.run(1)
    _ok(core, lib.cgfd_io_fault_end(fo, h))
    for o in so + [sn]:
        _ok(core, lib.cgfd_io_close(o))
    _ok(core, lib.cgfd_io_pg_write(h, d, None))
    fend = core.read_fault(0)

    names = ["Tn", "Ts1", "Ts2", "Vs", "Vs1", "Vs2", "Slip", "Slip1", "Slip2"]
    with netcdf_file(str(tmp_path / f"fault_i{i0 - 2}_px0_py0_pz0.nc"), "r", mmap=False) as f:
        assert f.dimensions["k"] == sc.nk and f.dimensions["j"] == sc.nj
        assert f.variables["time"].shape == (4,)
        assert np.allclose(f.variables["time"][:], np.arange(0, 20, 5) * dt)
        for q, n in enumerate(names):
            for r in range(4):
                assert np.array_equal(f.variables[n][r], faults[r][q]), (n, r)
        assert np.array_equal(f.variables["Peak_vs"][:], fend[9, 3:-3, 3:-3])
        assert np.array_equal(f.variables["Init_t0"][:], fend[10, 3:-3, 3:-3])
        assert int(f.i_index_with_ghosts_in_this_thread) == i0
        assert float(np.abs(f.variables["Slip"][3]).max()) > 0.0
        # the frames against the oracle's fault outputs at the same steps
        for q in (0, 3, 6):
            r = rel_l2(np.asarray(f.variables[names[q]][3], np.float64), o_faults[3][q])
            assert r <= 1e-3, (names[q], r)
        ofend = orc.read_fault(0)
        assert rel_l2(np.asarray(f.variables["Peak_vs"][:], np.float64), ofend[9, 3:-3, 3:-3]) <= 1e-3
    wn = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"]
    for s_, (ax, idx) in enumerate(slices):
        fn = ["slicex_i", "slicey_j", "slicez_k"][ax] + f"{idx - 2}_px0_py0_pz0.nc"
        with netcdf_file(str(tmp_path / fn), "r", mmap=False) as f:
            for q, n in enumerate(wn):
                for r in range(4):
                    assert np.array_equal(f.variables[n][r], slc[r][s_][q]), (fn, n, r)
            # the physical part of the last frame against the oracle's wavefield
            got = np.stack([np.asarray(f.variables[n][3], np.float64) for n in wn])
            assert rel_l2(got, o_slc[3][s_][:, 3:-3, 3:-3]) <= 1e-3, fn
    with netcdf_file(str(tmp_path / "snap_001_px0_py0_pz0.nc"), "r", mmap=False) as f:
        assert (f.dimensions["k"], f.dimensions["j"], f.dimensions["i"]) == (8, 7, 9)
        order = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy", "Exx", "Eyy", "Ezz", "Eyz", "Exz", "Exy"]
        for q, n in enumerate(order):
            for r in range(4):
                assert np.array_equal(f.variables[n][r], snaps[r][q]), (n, r)
        assert list(f.variables)[4:10] == ["Txx", "Tyy", "Tzz", "Txz", "Tyz", "Txy"]
        assert list(f.index_stride_in_this_thread) == [3, 5, 2]
    with netcdf_file(str(tmp_path / "PG_V_A_D_px0_py0_pz0.nc"), "r", mmap=False) as f:
        nz, ny, nx = sc.shape_ext
        assert f.dimensions["j"] == ny and f.dimensions["i"] == nx
        assert list(f.variables)[:3] == ["PGV", "PGVh", "PGVx"]
        buf = np.empty((15, ny, nx), np.float32)
        _ok(core, lib.cgfd_core_read_pg(h, buf.ctypes.data_as(C.c_void_p)))
        assert np.array_equal(f.variables["PGV"][:], buf[0])
        assert np.array_equal(f.variables["PGDz"][:], buf[14])
        assert float(buf[0].max()) > 0.0
        opg = orc.read_pg()
        assert rel_l2(np.asarray(f.variables["PGV"][:], np.float64), opg[0]) <= 1e-3
        assert rel_l2(np.asarray(f.variables["PGDz"][:], np.float64), opg[14]) <= 1e-3
    orc.close()
    core.close()
    st.close()


def _read_sac(path):
    raw = path.read_bytes()
    f = np.frombuffer(raw[:280], np.float32)
    n = np.frombuffer(raw[280:440], np.int32)
    k = raw[440:632]
    return f, n, k, np.frombuffer(raw[632:], np.float32)


def test_station_sac_files(built_libs, tmp_path):
    """SAC files of the station traces: header fields as the reference's
    sacExport1C1R sets them, samples equal to the recorded traces."""
    from cgfd import configs, model
    sc = configs.small(imethod=2)
    st = model.Setup(sc)
    core = model.Core(st)
    nt = 30
    core.set_stations(sc.stations + 3, nt)
    core.run(nt)
    tr = core.read_stations()
    ns = len(sc.stations)
    names = (C.c_char_p * ns)(*[f"ST{q:02d}".encode() for q in range(ns)])
    xyz = np.ascontiguousarray(np.arange(ns * 3, dtype=np.float32).reshape(ns, 3) * 100.0)
    _ok(core, core.lib.cgfd_io_stations_sac(core.handle, str(tmp_path).encode(), names, xyz.ctypes.data_as(C.c_void_p),
                                            1))
    dt = np.float32(st.desc.dt)
    for q, cmp in enumerate(["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"]):
        for s_ in (0, ns - 1):
            f, n, k, data = _read_sac(tmp_path / f"ST{s_:02d}.{cmp}.sac")
            assert np.array_equal(data, tr[s_, :, q])
            assert f[0] == dt and f[5] == 0 and f[6] == np.float32(dt * np.float32(nt - 1))
            assert list(f[31:35]) == [xyz[s_, 0], xyz[s_, 1], xyz[s_, 2], dt]
            assert list(f[35:39]) == [0, 0, 0, 0] and f[57] == 0 and f[58] == 0 and f[1] == -12345.0
            assert n[6] == 6 and n[9] == nt and n[15] == 1 and n[35] == 1 and n[0] == -12345
            assert k[:8] == b"-12345\0\0" and k[160:168] == b"-12345\0\0" and k[8:24] == b"-12345  " + bytes(8)
            assert k[24:32] == b"-12345  "
    # strain files: the reference's float/double mix, at the station node's media
    for s_ in (0, ns - 1):
        i, j, k_ = sc.stations[s_]
        lam, mu = np.float32(sc.lam[k_, j, i]), np.float32(sc.mu[k_, j, i])
        e1 = np.float32((np.float64(lam + mu)) / (np.float64(mu) * (3.0 * np.float64(lam) + 2.0 * np.float64(mu))))
        e2 = np.float32(-np.float64(lam) / (2.0 * np.float64(mu) * (3.0 * np.float64(lam) + 2.0 * np.float64(mu))))
        e3 = np.float32(1.0 / np.float64(mu))
        t = tr[s_].astype(np.float32)
        e0 = (e2 * (t[:, 3] + t[:, 4] + t[:, 5])).astype(np.float32)
        ref = {"Exx": e0 - (e2 - e1) * t[:, 3], "Eyy": e0 - (e2 - e1) * t[:, 4], "Ezz": e0 - (e2 - e1) * t[:, 5],
               "Eyz": (0.5 * np.float64(e3) * t[:, 6].astype(np.float64)).astype(np.float32),
               "Exz": (0.5 * np.float64(e3) * t[:, 7].astype(np.float64)).astype(np.float32),
               "Exy": (0.5 * np.float64(e3) * t[:, 8].astype(np.float64)).astype(np.float32)}
        for n_, r_ in ref.items():
            _, _, _, data = _read_sac(tmp_path / f"ST{s_:02d}.{n_}.sac")
            scale = max(float(np.abs(r_).max()), 1e-30)
            assert float(np.abs(data - r_).max()) <= 1e-6 * scale, n_
    core.close()
    st.close()
