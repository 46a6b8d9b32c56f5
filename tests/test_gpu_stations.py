"""Stations and fault receivers between grid nodes, through the C ABI.

* io_recv_keep (io_funcs.cu:427-469): an interpolated station records, at
  every step, the eight-term sum buff[t] Lx Ly Lz over the corners of its
  cell in the reference's term order and FP32 rounding, with corner 5 in
  the place of corner 4 (the reference's buff[5] twice, :459-460). The
  recorded traces are compared bit for bit with that sum evaluated in
  float32 from traces recorded at the eight corner nodes; with
  CGFD_COMPAT_FIX_STATION_INTERP the sum uses corner 4.
* io_fault_recv_keep (io_funcs.cu:2313-2351): the four-node bilinear sum on
  the fault plane, the same way.
* the adjoint of the interpolated recording: on a locked fault the traces
  are linear in the initial wavefield, and <g, w0> = sum r . T(w0).
"""
import numpy as np
import pytest

from helpers import progress

pytestmark = pytest.mark.gpu

NT = 24


def _reference_sum(corner_traces, d, q3):
    """corner_traces[t] for corners t = i + 2 j + 4 k, float32 (nt, 9)"""
    f32 = np.float32
    lx = (f32(1.0 - f32(d[0])), f32(d[0]))
    ly = (f32(1.0 - f32(d[1])), f32(d[1]))
    lz = (f32(1.0 - f32(d[2])), f32(d[2]))
    acc = None
    for u in range(8):
        src = 5 if (u == 4 and q3) else u
        term = ((corner_traces[src] * lx[u & 1]).astype(f32) * ly[(u >> 1) & 1]).astype(f32) * lz[u >> 2]
        term = term.astype(f32)
        acc = term if acc is None else (acc + term).astype(f32)
    return acc


@pytest.mark.parametrize("fix", [False, True])
def test_interpolated_stations_match_the_corner_sum(built_libs, fix):
    from cgfd import configs, lib as L, model
    sc = configs.small(imethod=2, curvilinear=True)
    st = model.Setup(sc)
    core = model.Core(st, compat_flags=L.COMPAT_FIX_STATION_INTERP if fix else 0)
    base = np.array([[12, 20, 9], [20, 25, 18], [15, 30, 28]], np.int32)   # ghost-inclusive
    d = np.array([[0.25, 0.5, 0.75], [0.1, 0.9, 0.3], [0.6, 0.2, 0.0]], np.float32)
    corners = np.array([[b[0] + (t & 1), b[1] + ((t >> 1) & 1), b[2] + (t >> 2)] for b in base for t in range(8)],
                       np.int32)
    progress(f"[interp stations fix={fix}] {len(base)} stations, {NT} steps")
    core.set_stations(base, NT, dijk=d)
    node = model.Core(st)
    node.set_stations(corners, NT)
    core.run(NT)
    node.run(NT)
    got = core.read_stations()
    nodes = node.read_stations().reshape(len(base), 8, NT, 9)
    for s in range(len(base)):
        ref = _reference_sum(nodes[s], d[s], q3=not fix)
        assert np.array_equal(got[s], ref), f"station {s}: max diff {np.abs(got[s] - ref).max():.3e}"
        assert np.abs(ref).max() > 0
    if not fix:
        # the corner-4 term really is corner 5's value (a visible difference)
        good = _reference_sum(nodes[0], d[0], q3=False)
        assert not np.array_equal(got[0], good)
    core.close()
    node.close()
    st.close()


def test_interpolated_fault_receivers_match_the_bilinear_sum(built_libs):
    from cgfd import configs, model
    sc = configs.small(imethod=2, curvilinear=True)
    st = model.Setup(sc)
    j1, j2, k1, k2 = sc.faults[0].grid
    jc, kc = (j1 + j2) // 2 + 3, sc.nk - 1 - 9 + 3
    base = np.array([[0, jc, kc], [0, jc - 3, kc + 2]], np.int32)
    d = np.array([[0.5, 0.25], [0.8, 0.6]], np.float32)
    nodes = np.array([[0, b[1] + (u & 1), b[2] + (u >> 1)] for b in base for u in range(4)], np.int32)
    a = model.Core(st)
    a.set_fault_receivers(base, NT, djk=d)
    b = model.Core(st)
    b.set_fault_receivers(nodes, NT)
    a.run(NT)
    b.run(NT)
    got = a.read_fault_receivers()
    nd = b.read_fault_receivers().reshape(len(base), 4, NT, 9)
    f32 = np.float32
    for r in range(len(base)):
        ly = (f32(1.0 - f32(d[r, 0])), f32(d[r, 0]))
        lz = (f32(1.0 - f32(d[r, 1])), f32(d[r, 1]))
        acc = None
        for u in range(4):
            t = ((nd[r, u] * ly[u & 1]).astype(f32) * lz[u >> 1]).astype(f32)
            acc = t if acc is None else (acc + t).astype(f32)
        assert np.array_equal(got[r], acc)
    assert np.abs(got[:, :, 3]).max() > 1e-3      # the receivers see slip
    a.close()
    b.close()
    st.close()


def test_interpolated_station_adjoint_identity(built_libs):
    """Locked fault, stations between nodes: the adjoint of the eight-term
    recording is its transpose."""
    from cgfd import model
    from test_gpu_adjoint_fault import _scenario, _smooth_field
    sc = _scenario(2, locked=True)
    st = model.Setup(sc)
    core = model.Core(st)
    rng = np.random.default_rng(31)
    base = (sc.stations + 3).astype(np.int32)
    base[:, 2] = np.minimum(base[:, 2], sc.nk + 1)   # room for k + 1 inside the grid
    d = rng.uniform(0.0, 1.0, (len(base), 3)).astype(np.float32)
    nt = 20
    core.set_stations(base, nt, dijk=d)
    core.adjoint_forward(nt)
    w0 = _smooth_field(sc, 7, 1e-3)
    core.write_wavefield(w0)
    core.run(nt)
    tr = core.read_stations()
    r = rng.standard_normal(tr.shape).astype(np.float32)
    r[:, :, 3:] *= 1e-6
    g, _ = core.adjoint_fault(r)
    lhs = float(np.sum(g.astype(np.float64) * w0.astype(np.float64)))
    rhs = float(np.sum(r.astype(np.float64) * tr.astype(np.float64)))
    print(f"[interp station adjoint] <g, w0> = {lhs:.9e}, sum r.T(w0) = {rhs:.9e}, rel {abs(lhs - rhs) / abs(rhs):.2e}",
          flush=True)
    assert abs(lhs - rhs) <= 1e-4 * abs(rhs)
    core.close()
    st.close()
