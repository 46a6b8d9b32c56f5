"""TEST INFRASTRUCTURE ONLY -- the CPU oracle.

ctypes wrapper of oracle/_build/libcgfd_oracle.so, the CPU restatement of
the reference time-stepping core (cgfd_oracle.cpp). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this package;
the product never does.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libcgfd_oracle.so"
_LIB = None


def load():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing; build with __graft_entry__.build()")
        lib = C.CDLL(str(LIB_PATH))
        vp, i32 = C.c_void_p, C.c_int
        sig = {
            "oracle_create": (vp, [vp, i32]),
            "oracle_destroy": (None, [vp]),
            "oracle_run": (i32, [vp, i32]),
            "oracle_set_threads": (i32, [i32]),
            "oracle_read_wave": (i32, [vp, vp]),
            "oracle_write_wave": (i32, [vp, vp]),
            "oracle_read_fault": (i32, [vp, i32, vp]),
            "oracle_read_fault_state": (i32, [vp, i32, vp]),
            "oracle_write_fault_state": (i32, [vp, i32, vp, vp]),
            "oracle_read_pg": (i32, [vp, vp]),
            "oracle_step_index": (i32, [vp]),
            "oracle_probe_inner": (i32, [vp, vp, i32, i32, i32, vp]),
            "oracle_read_freemats": (i32, [vp, vp]),
            "oracle_probe_friction": (i32, [vp, vp, i32]),
            "oracle_probe_w2f_f2w": (i32, [vp, C.c_double, vp, i32, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """CPU restatement of the core on a boundary descriptor (cgfd.lib.Desc)."""

    def __init__(self, desc, shape_ext, nfault, double=False, fix_q1=False, plant=0):
        """plant != 0 builds the oracle with a deliberate mistake (see
        oracle_create in cgfd_oracle.cpp); only the tests that show a pin
        catches that mistake use it."""
        self.lib = load()
        flags = (1 if double else 0) | (2 if fix_q1 else 0) | ((int(plant) & 0xFF) << 8)
        self.h = self.lib.oracle_create(C.byref(desc), flags)
        self.shape_ext = shape_ext
        self.nfault = nfault

    @classmethod
    def from_scenario(cls, sc, double=False, fix_q1=False, plant=0):
        """The oracle on its own FP64 setup of the scenario (setup_ref), not
        on the product's descriptor."""
        from .setup_ref import SetupRef
        ref = SetupRef(sc)
        o = cls(ref.desc(), sc.shape_ext, len(sc.faults), double=double, fix_q1=fix_q1, plant=plant)
        o.setup_ref = ref          # keeps the descriptor's arrays alive
        return o

    def run(self, n):
        self.lib.oracle_run(self.h, n)

    @staticmethod
    def threads(n=0):
        return load().oracle_set_threads(n)

    @property
    def step_index(self):
        return self.lib.oracle_step_index(self.h)

    def read_wavefield(self):
        nz, ny, nx = self.shape_ext
        out = np.empty((9, nz, ny, nx), np.float64)
        self.lib.oracle_read_wave(self.h, _p(out))
        return out

    def write_wavefield(self, w):
        w = np.ascontiguousarray(w, np.float64)
        self.lib.oracle_write_wave(self.h, _p(w))

    def read_fault(self, fid=0):
        nz, ny, _ = self.shape_ext
        out = np.empty((11, nz, ny), np.float64)
        assert self.lib.oracle_read_fault(self.h, fid, _p(out)) == 0
        return out

    def read_fault_state(self, fid=0):
        nz, ny, _ = self.shape_ext
        out = np.empty((9, 2, nz, ny), np.float64)
        assert self.lib.oracle_read_fault_state(self.h, fid, _p(out)) == 0
        return out

    def write_fault_state(self, state, fid=0, t1rec=None):
        nz, ny, _ = self.shape_ext
        st = np.ascontiguousarray(state, np.float64).reshape(9, 2, nz, ny)
        t1 = None if t1rec is None else np.ascontiguousarray(t1rec, np.float64).reshape(6, nz, ny)
        assert self.lib.oracle_write_fault_state(self.h, fid, _p(st), None if t1 is None else _p(t1)) == 0

    def read_pg(self):
        nz, ny, nx = self.shape_ext
        out = np.empty((15, ny, nx), np.float64)
        self.lib.oracle_read_pg(self.h, _p(out))
        return out

    def probe_inner(self, w, idir, jdir, kdir):
        w = np.ascontiguousarray(w, np.float64)
        out = np.zeros_like(w)
        self.lib.oracle_probe_inner(self.h, _p(w), idir, jdir, kdir, _p(out))
        return out

    def read_freemats(self):
        nz, ny, nx = self.shape_ext
        out = np.empty((2, ny, nx, 9), np.float64)
        self.lib.oracle_read_freemats(self.h, _p(out))
        return out

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def probe_friction(vals, double=True):
    v = np.ascontiguousarray(vals, np.float64)
    out = np.zeros(7, np.float64)
    load().oracle_probe_friction(_p(v), _p(out), 1 if double else 0)
    return out


def probe_w2f_f2w(metric9, jac, stress6, fix_q1):
    """wave2fault then fault2wave on one node with the oracle's own per-node
    steps; returns (stress written back [6] in {xx, yy, zz, yz, xz, xy}
    order, traction matrix [3][3])."""
    m = np.ascontiguousarray(metric9, np.float64)
    s = np.ascontiguousarray(stress6, np.float64)
    out = np.zeros(6, np.float64)
    tr = np.zeros(9, np.float64)
    load().oracle_probe_w2f_f2w(_p(m), float(jac), _p(s), int(bool(fix_q1)), _p(out), _p(tr))
    return out, tr.reshape(3, 3)
