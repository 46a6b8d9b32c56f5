"""Station lists on the CPU: the reference's station file (io_recv_read_locate,
io_funcs.cu:23-320) and fault station file (io_fault_recv_read_locate,
io_funcs.cu:2170-2300), read and located by cgfd_io_read_stations /
cgfd_io_read_fault_stations.

* grid-index stations: 1-based indices, truncation to the base node, the
  fractional rest as increments, depth indices counted down from the free
  surface, comments and empty lines skipped, stations off the grid dropped;
* coordinate stations on a curvilinear grid: a point placed by the trilinear
  map of a cell at a lattice position (multiples of 1/100 of the cell) is
  located to that cell and position; a grid node to itself with zero
  increments; a depth below the surface above a top-layer node becomes the
  node's elevation minus the depth;
* fault stations: base node and increments, an off-plane station is an
  error.
"""
import numpy as np
import pytest


def _setup(curvilinear=True):
    from cgfd import configs, model
    sc = configs.small(imethod=2, curvilinear=curvilinear)
    return sc, model.Setup(sc)


def _write(path, lines):
    path.write_text("\n".join(lines) + "\n")
    return path


def test_index_stations(built_libs, tmp_path):
    sc, st = _setup()
    nk = sc.nk
    f = _write(tmp_path / "station.list", [
        "# name flag_coord flag_depth x y z",
        "4",
        "",
        "a 0 0 5 6 7",                   # a node: (4, 5, 6) from 0
        "b 0 0 5.25 6.5 7.75",           # between nodes
        "c 0 1 10 12 2.5",               # depth index 2.5 below the surface
        f"d 0 0 {sc.ni + 5} 3 3",        # off the grid: dropped
    ])
    got = st.read_stations(f)
    assert [s["name"] for s in got] == ["a", "b", "c"]
    assert got[0]["ijk"] == (4 + 3, 5 + 3, 6 + 3) and got[0]["dijk"] == (0.0, 0.0, 0.0)
    assert got[1]["ijk"] == (4 + 3, 5 + 3, 6 + 3)
    assert np.allclose(got[1]["dijk"], (0.25, 0.5, 0.75), atol=1e-6)
    kz = (nk - 1) - 2.5
    assert got[2]["ijk"] == (9 + 3, 11 + 3, int(kz) + 3)
    assert np.isclose(got[2]["dijk"][2], kz - int(kz), atol=1e-6)
    xyz = st.coords()
    i, j, k = got[1]["ijk"]
    assert got[1]["xyz"] == tuple(float(xyz[c, k, j, i]) for c in range(3))
    st.close()


def _trilinear(xyz, i, j, k, a, b, c):
    """The point at lattice position (a, b, c) of cell (i, j, k), in float32
    as the locator's sampling evaluates it."""
    f32 = np.float32
    L = [(f32(1) - f32(a), f32(a)), (f32(1) - f32(b), f32(b)), (f32(1) - f32(c), f32(c))]
    out = []
    for comp in range(3):
        v = f32(0)
        for kk in range(2):
            for jj in range(2):
                for ii in range(2):
                    w = f32(f32(L[0][ii] * L[1][jj]) * L[2][kk])
                    v = f32(v + f32(w * xyz[comp, k + kk, j + jj, i + ii]))
        out.append(float(v))
    return out


def test_coordinate_stations_locate_to_their_cell(built_libs, tmp_path):
    sc, st = _setup(curvilinear=True)
    xyz = st.coords()
    cases = [((10, 12, 8), (0.25, 0.4, 0.75)), ((20, 30, 15), (0.6, 0.1, 0.3)), ((7, 40, 20), (0.0, 0.0, 0.0))]
    lines = [str(len(cases))]
    for n, ((i, j, k), (a, b, c)) in enumerate(cases):
        x, y, z = _trilinear(xyz, i, j, k, a, b, c)
        lines.append(f"s{n} 1 0 {x:.9g} {y:.9g} {z:.9g}")
    got = st.read_stations(_write(tmp_path / "coords.list", lines))
    assert len(got) == len(cases)
    for s, ((i, j, k), (a, b, c)) in zip(got, cases):
        assert s["ijk"] == (i, j, k), (s, i, j, k)
        assert np.allclose(s["dijk"], (a, b, c), atol=2e-6), s
    st.close()


def test_coordinate_stations_by_depth(built_libs, tmp_path):
    """flag_depth 1: z is a depth below the free surface at (x, y); above a
    top-layer node the surface elevation is the node's own."""
    sc, st = _setup(curvilinear=True)
    xyz = st.coords()
    ktop = sc.nk + 2
    i, j = 15, 21
    x, y, zt = (float(xyz[c, ktop, j, i]) for c in range(3))
    depth = float(zt - xyz[2, ktop - 6, j, i])      # exactly six layers down in this column
    got = st.read_stations(_write(tmp_path / "depth.list", ["1", f"d 1 1 {x:.9g} {y:.9g} {depth:.9g}"]))
    assert len(got) == 1
    assert np.isclose(got[0]["xyz"][2], np.float32(zt) - np.float32(depth), rtol=0, atol=1e-3)
    st.close()


def test_fault_stations(built_libs, tmp_path):
    sc, st = _setup()
    f = _write(tmp_path / "fault_station.list", ["3", "r1 1 20 15", "r2 1 20.5 15.25", "# end"])
    with pytest.raises(RuntimeError):
        st.read_fault_stations(f)                   # the count says 3, the file holds 2
    f = _write(tmp_path / "fault_station.list", ["2", "r1 1 20 15", "r2 1 20.5 15.25"])
    got = st.read_fault_stations(f)
    assert [s["name"] for s in got] == ["r1", "r2"]
    assert got[0]["jk"] == (19 + 3, 14 + 3) and got[0]["djk"] == (0.0, 0.0)
    assert got[1]["jk"] == (19 + 3, 14 + 3) and np.allclose(got[1]["djk"], (0.5, 0.25))
    xyz = st.coords()
    i0 = sc.faults[0].x_index + 3
    assert got[0]["xyz"] == tuple(float(xyz[c, 17, 22, i0]) for c in range(3))
    bad = _write(tmp_path / "bad.list", ["1", f"r 1 {sc.nj + 2} 3"])
    with pytest.raises(RuntimeError, match="outside the calculation area"):
        st.read_fault_stations(bad)
    st.close()


def test_receiver_line(built_libs):
    """A receiver line (io_line_locate, io_funcs.cu:312-425): nodes start + n
    incre inside the grid, named <line>.no<n> as the reference names their
    SAC files, with their coordinates; nodes past the grid are skipped."""
    sc, st = _setup()
    got = st.line_stations("line_x", (2, 10, sc.nk - 1), (5, 0, 0), 10)
    n_in = len([n for n in range(10) if 2 + 5 * n <= sc.ni - 1])
    assert len(got) == n_in and n_in < 10
    assert [s["name"] for s in got] == [f"line_x.no{n}" for n in range(n_in)]
    xyz = st.coords()
    for n, s in enumerate(got):
        assert s["ijk"] == (2 + 5 * n + 3, 13, sc.nk + 2) and s["dijk"] == (0.0, 0.0, 0.0)
        assert s["xyz"] == tuple(float(xyz[c, sc.nk + 2, 13, 2 + 5 * n + 3]) for c in range(3))
    st.close()


def test_receiver_line_keeps_the_point_numbers(built_libs):
    """io_line_locate numbers a kept node by its place on the line
    (recv_seq = ipt, io_funcs.cu:412), skipped nodes included."""
    sc, st = _setup()
    got = st.line_stations("l2", (-8, 4, 5), (5, 1, 0), 4)
    assert [s["name"] for s in got] == ["l2.no2", "l2.no3"]
    assert [s["ijk"] for s in got] == [(2 + 3, 6 + 3, 5 + 3), (7 + 3, 7 + 3, 5 + 3)]
    assert st.line_stations("empty", (0, 0, 0), (1, 1, 1), 0) == []
    st.close()
