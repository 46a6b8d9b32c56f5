"""Scenario -> host setup -> boundary descriptor -> GPU core.

A :class:`Scenario` holds what the reference reads from its par file and
input NetCDF files (grid, media, boundaries, faults and their initial
stress/friction, time stepping). :class:`Setup` runs the C++ host setup of
the driver on it (include/cgfd_setup.h) and exposes the boundary descriptor;
:class:`Core` drives the HIP time-stepping core through the C ABI.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import lib as L

FAULT_FIELDS = ("Tx", "Ty", "Tz", "mu_s", "mu_d", "Dc", "C0")  # init_stress_<id>.nc variables


@dataclass
class FaultSpec:
    x_index: int                      # 0-based physical x index of the plane
    grid: tuple                       # 0-based (j1, j2, k1, k2) = fault_grid
    fields: dict                      # FAULT_FIELDS -> float32 (nk, nj)


@dataclass
class Scenario:
    ni: int
    nj: int
    nk: int
    x: np.ndarray                     # (nk, nj, ni) float32 physical coordinates
    y: np.ndarray
    z: np.ndarray
    lam: np.ndarray                   # (nk, nj, ni) float32
    mu: np.ndarray
    rho: np.ndarray                   # density
    free_top: bool = True
    pml_layers: list = field(default_factory=lambda: [[0, 0], [0, 0], [0, 0]])
    pml_alpha_max: float = 3.14
    pml_beta_max: float = 2.0
    pml_velocity: float = 7000.0
    imethod: int = 2
    faults: list = field(default_factory=list)
    united_margin: int = 40
    dt: float = 0.008
    t0: float = 0.0
    nt_total: int = 1
    stations: np.ndarray | None = None   # (n, 3) physical (i, j, k)
    name: str = "scenario"

    @property
    def shape_ext(self):
        return (self.nk + 6, self.nj + 6, self.ni + 6)


class Setup:
    """Host setup (metric, PML profiles, fault coefficients and masks)."""

    def __init__(self, sc: Scenario, device_id: int = 0, device_setup: bool = False, metric=None):
        """device_setup: compute the metric and the fault coefficients with
        the HIP kernels of the setup (same arithmetic as the host loops).
        metric: an imported metric, (10, nk, nj, ni) at the physical points
        in the order jac, xi_x..xi_z, eta_x..eta_z, zeta_x..zeta_z, used
        instead of the one the coordinates give."""
        self.lib = L.load_library()
        self.sc = sc
        keep = []

        def arr(a, dtype=np.float32):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return L.ptr(a)

        p = L.SetupParams()
        p.ni, p.nj, p.nk = sc.ni, sc.nj, sc.nk
        p.x, p.y, p.z = arr(sc.x), arr(sc.y), arr(sc.z)
        p.lambda_, p.mu, p.rho = arr(sc.lam), arr(sc.mu), arr(sc.rho)
        p.free_top = int(sc.free_top)
        for d in range(3):
            for s in range(2):
                p.pml_layers[d][s] = int(sc.pml_layers[d][s])
                p.pml_alpha_max[d][s] = sc.pml_alpha_max
                p.pml_beta_max[d][s] = sc.pml_beta_max
                p.pml_velocity[d][s] = sc.pml_velocity
        p.imethod = sc.imethod
        p.number_fault = len(sc.faults)
        for fid, f in enumerate(sc.faults):
            p.fault_x_index[fid] = f.x_index
            for q in range(4):
                p.fault_grid[4 * fid + q] = int(f.grid[q])
            for q, name in enumerate(FAULT_FIELDS):
                p.init_stress[fid][q] = arr(f.fields[name])
        p.united_margin = sc.united_margin
        p.dt, p.t0, p.nt_total, p.device_id = sc.dt, sc.t0, sc.nt_total, device_id
        if metric is not None:
            m = np.asarray(metric)
            if m.shape != (10, sc.nk, sc.nj, sc.ni):
                raise ValueError(f"metric must be (10, nk, nj, ni) = {(10, sc.nk, sc.nj, sc.ni)}, got {m.shape}")
            p.metric = arr(m)
        p.device_setup = int(bool(device_setup))
        h = C.c_void_p()
        L.check(self.lib, self.lib.cgfd_setup_create(C.byref(p), C.byref(h)))
        del keep  # the setup copied everything it needs
        self.handle = h
        self.desc = L.Desc()
        L.check(self.lib, self.lib.cgfd_setup_get_desc(h, C.byref(self.desc)))

    def read_stations(self, path) -> list:
        """The station list file at path, located on this grid
        (cgfd_io_read_stations): dicts with name, ijk (ghost-inclusive base
        node), dijk (increments) and xyz, in file order, stations off the
        grid dropped."""
        n_file, n_loc = C.c_int(), C.c_int()
        L.check(self.lib, self.lib.cgfd_io_read_stations(str(path).encode(), self.handle, None, 0, C.byref(n_file),
                                                         C.byref(n_loc)))
        arr = (L.Station * max(n_file.value, 1))()
        L.check(self.lib, self.lib.cgfd_io_read_stations(str(path).encode(), self.handle, arr, n_file.value,
                                                         C.byref(n_file), C.byref(n_loc)))
        return [{"name": s.name.decode(), "ijk": (s.i, s.j, s.k), "dijk": (s.di, s.dj, s.dk), "xyz": (s.x, s.y, s.z)}
                for s in arr[:n_loc.value]]

    def line_stations(self, name: str, start, incre, count: int) -> list:
        """A receiver line (cgfd_io_line_stations): start, incre = global
        physical (i, j, k) from 0; the nodes inside the grid as station
        dicts named <name>.no<n>."""
        st = np.ascontiguousarray(start, np.int32)
        inc = np.ascontiguousarray(incre, np.int32)
        arr = (L.Station * max(count, 1))()
        n = C.c_int()
        L.check(self.lib, self.lib.cgfd_io_line_stations(self.handle, name.encode(), L.ptr(st), L.ptr(inc), count, arr,
                                                         C.byref(n)))
        return [{"name": s.name.decode(), "ijk": (s.i, s.j, s.k), "dijk": (s.di, s.dj, s.dk), "xyz": (s.x, s.y, s.z)}
                for s in arr[:n.value]]

    def read_fault_stations(self, path) -> list:
        """The fault station file at path (cgfd_io_read_fault_stations):
        dicts with name, fault_id, jk (ghost-inclusive base node), djk and
        xyz."""
        n = C.c_int()
        L.check(self.lib, self.lib.cgfd_io_read_fault_stations(str(path).encode(), self.handle, None, 0,
                                                               C.byref(n)))
        arr = (L.FaultStation * max(n.value, 1))()
        L.check(self.lib, self.lib.cgfd_io_read_fault_stations(str(path).encode(), self.handle, arr, n.value,
                                                               C.byref(n)))
        return [{"name": s.name.decode(), "fault_id": s.fault_id, "jk": (s.j, s.k), "djk": (s.dj, s.dk),
                 "xyz": (s.x, s.y, s.z)} for s in arr[:n.value]]

    def dt_max(self) -> float:
        v = C.c_float()
        L.check(self.lib, self.lib.cgfd_setup_dt_max(self.handle, C.byref(v)))
        return float(v.value)

    def coords(self) -> np.ndarray:
        nz, ny, nx = self.sc.shape_ext
        n = 3 * nz * ny * nx
        pp = self.lib.cgfd_setup_coords(self.handle)
        return np.ctypeslib.as_array(pp, shape=(n,)).reshape(3, nz, ny, nx).copy()

    def metric(self) -> np.ndarray:
        nz, ny, nx = self.sc.shape_ext
        pp = C.cast(self.desc.metric, C.POINTER(C.c_float))
        return np.ctypeslib.as_array(pp, shape=(10 * nz * ny * nx,)).reshape(10, nz, ny, nx).copy()

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cgfd_setup_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _autotune(obj, reps):
    v, z, ms = C.c_int(), C.c_int(), C.c_float()
    L.check(obj.lib, obj.lib.cgfd_core_autotune(obj.handle, reps, C.byref(v), C.byref(z), C.byref(ms)))
    return v.value, z.value, ms.value


def _variant_info(obj):
    v, z = C.c_int(), C.c_int()
    L.check(obj.lib, obj.lib.cgfd_core_get_variant(obj.handle, C.byref(v), C.byref(z)))
    return v.value, z.value


class Core:
    """The GPU time-stepping core (one HIP stream, one device)."""

    def __init__(self, setup: Setup, compat_flags: int = 0, nan_every: int = 0, variant: int = 0):
        self.lib = L.load_library()
        self.setup = setup
        d = L.Desc.from_buffer_copy(setup.desc)
        d.compat_flags = compat_flags
        d.qc_check_nan_every = nan_every
        self.desc = d
        h = C.c_void_p()
        L.check(self.lib, self.lib.cgfd_core_create(C.byref(d), C.byref(h)))
        self.handle = h
        if variant:
            L.check(self.lib, self.lib.cgfd_core_set_variant(h, variant))
        sc = setup.sc
        self.shape_ext = sc.shape_ext
        self.nfault = len(sc.faults)

    def set_variant(self, v: int):
        L.check(self.lib, self.lib.cgfd_core_set_variant(self.handle, v))

    def set_graph(self, on: bool):
        L.check(self.lib, self.lib.cgfd_core_set_graph(self.handle, int(on)))

    def set_overlap(self, on: bool):
        L.check(self.lib, self.lib.cgfd_core_set_overlap(self.handle, int(on)))

    def adjoint_initial(self, residual: np.ndarray) -> np.ndarray:
        """dJ/dw at the first recorded step for J = sum residual . traces
        (residual shaped like read_stations(): [nsta][nt][9])."""
        residual = np.ascontiguousarray(residual, np.float32)
        nsta, nt, nc = residual.shape
        if nc != 9:
            raise ValueError("residual must be [nsta][nt][9]")
        nz, ny, nx = self.shape_ext
        out = np.zeros((9, nz, ny, nx), np.float32)
        L.check(self.lib, self.lib.cgfd_adj_backward(self.handle, L.ptr(residual), nt, L.ptr(out)))
        return out

    def adjoint_forward(self, nt: int):
        """Arms the friction tape of the adjoint for the next nt steps (call it
        at the step the stations were set, before running them)."""
        L.check(self.lib, self.lib.cgfd_adj_forward(self.handle, int(nt)))

    def adjoint_fault(self, residual: np.ndarray, residual_fault: np.ndarray | None = None):
        """(dJ/dw at the first recorded step, dJ/d(fault parameters)) for
        J = sum residual . traces + sum residual_fault . fault-receiver
        records; the second is [nfault][7][nk][nj] over (mu_s, mu_d, Dc, C0,
        T0x, T0y, T0z)."""
        residual = np.ascontiguousarray(residual, np.float32)
        nsta, nt, nc = residual.shape
        if nc != 9:
            raise ValueError("residual must be [nsta][nt][9]")
        rf = None
        if residual_fault is not None:
            rf = np.ascontiguousarray(residual_fault, np.float32)
            if rf.ndim != 3 or rf.shape[1] != nt or rf.shape[2] != 9:
                raise ValueError("residual_fault must be [nrec][nt][9]")
        nz, ny, nx = self.shape_ext
        out = np.zeros((9, nz, ny, nx), np.float32)
        sc = self.setup.sc
        gf = np.zeros((len(sc.faults), 7, sc.nk, sc.nj), np.float32)
        L.check(self.lib, self.lib.cgfd_adj_backward_fault(self.handle, L.ptr(residual), L.ptr(rf) if rf is not None
                                                           else None, nt, L.ptr(out), L.ptr(gf)))
        return out, gf

    def read_slice(self, axis: int, index_ext: int) -> np.ndarray:
        """The nine components on the plane at ghost-inclusive `index_ext`
        along axis 0 (x: [9][nk][nj]), 1 (y: [9][nk][ni]) or 2 (z: [9][nj][ni])."""
        sc = self.setup.sc
        shape = {0: (9, sc.nk, sc.nj), 1: (9, sc.nk, sc.ni), 2: (9, sc.nj, sc.ni)}[axis]
        out = np.empty(shape, np.float32)
        L.check(self.lib, self.lib.cgfd_core_read_slice(self.handle, axis, index_ext, L.ptr(out)))
        return out

    def read_snapshot(self, box, flags: int = 3) -> np.ndarray:
        """box = (i1, ni, di, j1, nj, dj, k1, nk, dk), ghost-inclusive starts;
        flags: 1 velocity, 2 stress, 4 strain; [nfield][nk][nj][ni]."""
        b = np.ascontiguousarray(box, np.int32)
        nf = (3 if flags & 1 else 0) + (6 if flags & 2 else 0) + (6 if flags & 4 else 0)
        out = np.empty((nf, int(b[7]), int(b[4]), int(b[1])), np.float32)
        L.check(self.lib, self.lib.cgfd_core_read_snapshot(self.handle, L.ptr(b), flags, L.ptr(out)))
        return out

    def set_fault_receivers(self, fjk_ext: np.ndarray, nt_max: int, djk=None):
        """Receivers at (fault id, ghost-inclusive j, k) nodes of the fault
        planes, recording Tn..Slip2 at every step start; djk (n, 2):
        increments towards j+1, k+1 (bilinear between the nodes)."""
        a = np.ascontiguousarray(fjk_ext, np.int32).reshape(-1, 3)
        if djk is None:
            L.check(self.lib, self.lib.cgfd_core_set_fault_receivers(self.handle, a.shape[0], L.ptr(a), nt_max))
        else:
            d = np.ascontiguousarray(djk, np.float32).reshape(-1, 2)
            L.check(self.lib, self.lib.cgfd_core_set_fault_receivers_interp(self.handle, a.shape[0], L.ptr(a),
                                                                             L.ptr(d), nt_max))
        self._nfr = a.shape[0]
        self._nfr_nt = nt_max

    def read_fault_receivers(self) -> np.ndarray:
        out = np.zeros((self._nfr, self._nfr_nt, 9), np.float32)
        nt = C.c_int(0)
        L.check(self.lib, self.lib.cgfd_core_read_fault_receivers(self.handle, L.ptr(out), C.byref(nt)))
        flat = out.reshape(-1)[: self._nfr * nt.value * 9]
        return flat.reshape(self._nfr, nt.value, 9).copy()

    def state_bytes(self) -> int:
        n = C.c_size_t()
        L.check(self.lib, self.lib.cgfd_core_state_bytes(self.handle, C.byref(n)))
        return int(n.value)

    def save_state(self, dev_ptr: int = 0) -> np.ndarray | None:
        """Step-boundary state into a new host array, or into device memory
        at dev_ptr (state_bytes() bytes) when given."""
        if dev_ptr:
            L.check(self.lib, self.lib.cgfd_core_save_state(self.handle, C.c_void_p(dev_ptr), 1))
            return None
        out = np.empty(self.state_bytes(), np.uint8)
        L.check(self.lib, self.lib.cgfd_core_save_state(self.handle, L.ptr(out), 0))
        return out

    def load_state(self, state: np.ndarray | None = None, dev_ptr: int = 0):
        if dev_ptr:
            L.check(self.lib, self.lib.cgfd_core_load_state(self.handle, C.c_void_p(dev_ptr), 1))
            return
        state = np.ascontiguousarray(state, np.uint8)
        if state.size != self.state_bytes():
            raise ValueError(f"state has {state.size} bytes, this core needs {self.state_bytes()}")
        L.check(self.lib, self.lib.cgfd_core_load_state(self.handle, L.ptr(state), 0))

    def set_zchunk(self, z: int):
        L.check(self.lib, self.lib.cgfd_core_set_zchunk(self.handle, z))

    def set_vertical_columns(self, on: bool):
        """on: the five-term metric form on grids of vertical columns (the
        default); off: the nine-term form everywhere."""
        L.check(self.lib, self.lib.cgfd_core_set_vertical_columns(self.handle, int(on)))

    def vertical_columns(self):
        """(the grid has the vertical-column form, the volume update uses it)"""
        g, u = C.c_int(), C.c_int()
        L.check(self.lib, self.lib.cgfd_core_get_vertical_columns(self.handle, C.byref(g), C.byref(u)))
        return bool(g.value), bool(u.value)

    def autotune(self, reps: int = 3):
        """Pick the fastest volume-kernel variant and z chunk for this grid
        (recorded for every later core of the same shape)."""
        return _autotune(self, reps)

    def variant_info(self):
        return _variant_info(self)

    def run(self, nsteps: int):
        L.check(self.lib, self.lib.cgfd_core_run(self.handle, nsteps))

    def run_timed(self, nsteps: int):
        tot, vol, nv = C.c_float(), C.c_float(), C.c_int()
        L.check(self.lib, self.lib.cgfd_core_run_timed(self.handle, nsteps, C.byref(tot), C.byref(vol), C.byref(nv)))
        return tot.value, vol.value, nv.value

    def sync(self):
        L.check(self.lib, self.lib.cgfd_core_sync(self.handle))

    @property
    def step_index(self) -> int:
        return self.lib.cgfd_core_step_index(self.handle)

    def read_wavefield(self) -> np.ndarray:
        nz, ny, nx = self.shape_ext
        out = np.empty((9, nz, ny, nx), np.float32)
        L.check(self.lib, self.lib.cgfd_core_read_wavefield(self.handle, L.ptr(out)))
        return out

    def write_wavefield(self, w: np.ndarray):
        w = np.ascontiguousarray(w, np.float32)
        L.check(self.lib, self.lib.cgfd_core_write_wavefield(self.handle, L.ptr(w)))

    def read_fault(self, fid: int = 0) -> np.ndarray:
        nz, ny, _ = self.shape_ext
        out = np.empty((11, nz, ny), np.float32)
        L.check(self.lib, self.lib.cgfd_core_read_fault(self.handle, fid, L.ptr(out)))
        return out

    def read_fault_state(self, fid: int = 0) -> np.ndarray:
        nz, ny, _ = self.shape_ext
        out = np.empty((9, 2, nz, ny), np.float32)
        L.check(self.lib, self.lib.cgfd_core_read_fault_state(self.handle, fid, L.ptr(out)))
        return out

    def write_fault_state(self, state: np.ndarray, fid: int = 0, t1rec: np.ndarray | None = None):
        """Overwrites the split-node state [9][2][nz][ny]; t1rec [6][nz][ny]
        (hT1 x,y,z, mT1 x,y,z) optionally sets method 2's T1 recurrence."""
        nz, ny, _ = self.shape_ext
        st = np.ascontiguousarray(state, np.float32)
        if st.shape != (9, 2, nz, ny):
            raise ValueError(f"state must be (9, 2, {nz}, {ny})")
        t1 = None
        if t1rec is not None:
            t1 = np.ascontiguousarray(t1rec, np.float32)
            if t1.shape != (6, nz, ny):
                raise ValueError(f"t1rec must be (6, {nz}, {ny})")
        L.check(self.lib, self.lib.cgfd_core_write_fault_state(self.handle, fid, L.ptr(st), L.ptr(t1)))

    def read_pg(self) -> np.ndarray:
        nz, ny, nx = self.shape_ext
        out = np.empty((15, ny, nx), np.float32)
        L.check(self.lib, self.lib.cgfd_core_read_pg(self.handle, L.ptr(out)))
        return out

    def set_stations(self, ijk_ext: np.ndarray, nt_max: int, dijk=None):
        """Stations at ghost-inclusive (i, j, k) nodes; dijk (n, 3):
        increments towards i+1, j+1, k+1 (the reference's eight-point
        interpolation, cgfd_core_set_stations_interp)."""
        a = np.ascontiguousarray(ijk_ext, np.int32).reshape(-1, 3)
        if dijk is None:
            L.check(self.lib, self.lib.cgfd_core_set_stations(self.handle, a.shape[0], L.ptr(a), nt_max))
        else:
            d = np.ascontiguousarray(dijk, np.float32).reshape(-1, 3)
            L.check(self.lib, self.lib.cgfd_core_set_stations_interp(self.handle, a.shape[0], L.ptr(a), L.ptr(d),
                                                                      nt_max))
        self._nst = a.shape[0]
        self._ntmax = nt_max

    def read_stations(self) -> np.ndarray:
        buf = np.empty((self._nst, self._ntmax, 9), np.float32)
        nt = C.c_int()
        L.check(self.lib, self.lib.cgfd_core_read_stations(self.handle, L.ptr(buf), C.byref(nt)))
        n = nt.value
        flat = buf.reshape(-1)[: self._nst * n * 9]
        return flat.reshape(self._nst, n, 9).copy()

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cgfd_core_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several rupture scenarios on one grid (include/cgfd_core.h, batched
    ensembles). ``scenarios`` is a list of (faults, nt) where ``faults`` is a
    list of FaultSpec-like field dicts (one per fault of the setup, keys
    FAULT_FIELDS) and nt the scenario's number of steps; the first scenario
    may be None to reuse the setup's own fault inputs."""

    def __init__(self, setup: Setup, scenarios, variant: int = 0):
        self.lib = L.load_library()
        self.setup = setup
        nf = len(setup.sc.faults)
        descs = (L.Desc * len(scenarios))()
        self._keep = []
        for s, (fields, nt) in enumerate(scenarios):
            if fields is None:
                d = L.Desc.from_buffer_copy(setup.desc)
                d.nt_total = nt
            else:
                arrs = []
                for fid in range(nf):
                    for name in FAULT_FIELDS:
                        a = np.ascontiguousarray(fields[fid][name], np.float32)
                        arrs.append(a)
                self._keep.append(arrs)
                ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
                d = L.Desc()
                L.check(self.lib, self.lib.cgfd_setup_scenario_desc(setup.handle, ptrs, nt, C.byref(d)))
            descs[s] = d
        h = C.c_void_p()
        L.check(self.lib, self.lib.cgfd_batch_create(descs, len(scenarios), C.byref(h)))
        self.handle = h
        self.n = len(scenarios)
        if variant:
            L.check(self.lib, self.lib.cgfd_core_set_variant(h, variant))
        self.shape_ext = setup.sc.shape_ext

    def run(self, nsteps: int):
        L.check(self.lib, self.lib.cgfd_core_run(self.handle, nsteps))

    def sync(self):
        L.check(self.lib, self.lib.cgfd_core_sync(self.handle))

    def set_graph(self, on: bool):
        L.check(self.lib, self.lib.cgfd_core_set_graph(self.handle, int(on)))

    def set_overlap(self, on: bool):
        L.check(self.lib, self.lib.cgfd_core_set_overlap(self.handle, int(on)))

    def set_variant(self, v: int):
        L.check(self.lib, self.lib.cgfd_core_set_variant(self.handle, v))

    def set_zchunk(self, z: int):
        L.check(self.lib, self.lib.cgfd_core_set_zchunk(self.handle, z))

    def set_vertical_columns(self, on: bool):
        """on: the five-term metric form on grids of vertical columns (the
        default); off: the nine-term form everywhere."""
        L.check(self.lib, self.lib.cgfd_core_set_vertical_columns(self.handle, int(on)))

    def vertical_columns(self):
        """(the grid has the vertical-column form, the volume update uses it)"""
        g, u = C.c_int(), C.c_int()
        L.check(self.lib, self.lib.cgfd_core_get_vertical_columns(self.handle, C.byref(g), C.byref(u)))
        return bool(g.value), bool(u.value)

    def autotune(self, reps: int = 3):
        return _autotune(self, reps)

    def variant_info(self):
        return _variant_info(self)

    @property
    def active(self) -> int:
        return self.lib.cgfd_batch_active(self.handle)

    def steps(self, s: int) -> int:
        return self.lib.cgfd_batch_steps(self.handle, s)

    def read_wavefield(self, s: int) -> np.ndarray:
        nz, ny, nx = self.shape_ext
        out = np.empty((9, nz, ny, nx), np.float32)
        L.check(self.lib, self.lib.cgfd_batch_read_wavefield(self.handle, s, L.ptr(out)))
        return out

    def read_fault(self, s: int, fid: int = 0) -> np.ndarray:
        nz, ny, _ = self.shape_ext
        out = np.empty((11, nz, ny), np.float32)
        L.check(self.lib, self.lib.cgfd_batch_read_fault(self.handle, s, fid, L.ptr(out)))
        return out

    def read_pg(self, s: int) -> np.ndarray:
        _, ny, nx = self.shape_ext
        out = np.empty((15, ny, nx), np.float32)
        L.check(self.lib, self.lib.cgfd_batch_read_pg(self.handle, s, L.ptr(out)))
        return out

    def set_stations(self, ijk_ext: np.ndarray, nt_max: int):
        a = np.ascontiguousarray(ijk_ext, np.int32).reshape(-1, 3)
        L.check(self.lib, self.lib.cgfd_core_set_stations(self.handle, a.shape[0], L.ptr(a), nt_max))
        self._nst, self._ntmax = a.shape[0], nt_max

    def read_stations(self, s: int) -> np.ndarray:
        buf = np.empty((self._nst, self._ntmax, 9), np.float32)
        nt = C.c_int()
        L.check(self.lib, self.lib.cgfd_batch_read_stations(self.handle, s, L.ptr(buf), C.byref(nt)))
        flat = buf.reshape(-1)[: self._nst * nt.value * 9]
        return flat.reshape(self._nst, nt.value, 9).copy()

    def set_fault_receivers(self, fjk_ext: np.ndarray, nt_max: int):
        """Receivers at (fault id, ghost-inclusive j, k) nodes, recorded for
        every active scenario (cgfd_core_set_fault_receivers)."""
        a = np.ascontiguousarray(fjk_ext, np.int32).reshape(-1, 3)
        L.check(self.lib, self.lib.cgfd_core_set_fault_receivers(self.handle, a.shape[0], L.ptr(a), nt_max))
        self._nfr, self._nfr_nt = a.shape[0], nt_max

    def read_fault_receivers(self, s: int) -> np.ndarray:
        """Scenario s's receiver records [nrec][steps it recorded][9]."""
        buf = np.zeros((self._nfr, self._nfr_nt, 9), np.float32)
        nt = C.c_int()
        L.check(self.lib, self.lib.cgfd_batch_read_fault_receivers(self.handle, s, L.ptr(buf), C.byref(nt)))
        flat = buf.reshape(-1)[: self._nfr * nt.value * 9]
        return flat.reshape(self._nfr, nt.value, 9).copy()

    def adjoint_forward(self, nt: int):
        """Arms every scenario's friction tape for the next nt steps."""
        L.check(self.lib, self.lib.cgfd_adj_forward(self.handle, int(nt)))

    def adjoint_fault(self, s: int, residual: np.ndarray):
        """(dJ/dw at the first recorded step, dJ/d(fault parameters)) of
        scenario s for J = sum residual . its station traces."""
        residual = np.ascontiguousarray(residual, np.float32)
        nsta, nt, _ = residual.shape
        nz, ny, nx = self.shape_ext
        out = np.zeros((9, nz, ny, nx), np.float32)
        sc = self.setup.sc
        gf = np.zeros((len(sc.faults), 7, sc.nk, sc.nj), np.float32)
        L.check(self.lib, self.lib.cgfd_batch_adj_backward(self.handle, s, L.ptr(residual), nt, L.ptr(out), L.ptr(gf)))
        return out, gf

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cgfd_core_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
