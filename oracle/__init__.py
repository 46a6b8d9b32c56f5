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
# the same restatement built -O3 for AVX2/FMA, timed as the CPU baseline only
FAST_LIB_PATH = Path(__file__).resolve().parent / "_build" / "libcgfd_oracle_fast.so"
_LIB = None
_LIB_FILE = None


def _cpu_has_avx2_fma() -> bool:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("flags"):
                f = set(line.split(":", 1)[1].split())
                return {"avx2", "fma", "bmi2"} <= f
    except OSError:
        pass
    return False


def load_fast() -> bool:
    """Loads the -O3 AVX2/FMA build for timing if nothing is loaded yet and
    this CPU can run it; returns whether it is the loaded build."""
    if _LIB is None and FAST_LIB_PATH.exists() and _cpu_has_avx2_fma():
        load(FAST_LIB_PATH)
    return _LIB_FILE == FAST_LIB_PATH


def active_build() -> str:
    if _LIB_FILE == FAST_LIB_PATH:
        return "g++ -O3 -march=x86-64-v3 -fopenmp"
    return "g++ -O2 -march=x86-64-v2 -ffp-contract=off -fopenmp"


def load(path=None):
    global _LIB, _LIB_FILE
    if _LIB is None:
        path = Path(path) if path is not None else LIB_PATH
        if not path.exists():
            raise RuntimeError(f"{path} missing; build with __graft_entry__.build()")
        lib = C.CDLL(str(path))
        _LIB_FILE = path
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
            "oracle_read_kink_margin": (i32, [vp, i32, vp]),
            "oracle_write_fault_state": (i32, [vp, i32, vp, vp]),
            "oracle_read_pg": (i32, [vp, vp]),
            "oracle_step_index": (i32, [vp]),
            "oracle_probe_inner": (i32, [vp, vp, i32, i32, i32, vp]),
            "oracle_read_freemats": (i32, [vp, vp]),
            "oracle_probe_friction": (i32, [vp, vp, i32]),
            "oracle_probe_w2f_f2w": (i32, [vp, C.c_double, vp, i32, vp, vp]),
            "oracle_ad_gradient": (i32, [vp, i32, vp, i32, i32, vp, vp, i32, vp, vp, vp, vp, vp]),
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

    def read_kink_margin(self, fid=0):
        """(4, nz, ny): the smallest relative distance of each friction
        branch decision (shear vs strength, slip vs Dc, normal traction vs 0,
        |Ts| vs the 0.1 guard) from its switch point over all trial
        evaluations so far; 1e300 where the trial never ran."""
        nz, ny, _ = self.shape_ext
        out = np.empty((nz, ny, 4), np.float64)
        assert self.lib.oracle_read_kink_margin(self.h, fid, _p(out)) == 0
        return np.moveaxis(out, -1, 0).copy()

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


def ad_gradient(desc, shape_ext, nfault, w0, nt, sta_ijk, res, rec_fjk=None, res_f=None, fix_q1=False):
    """FP64 reverse-mode gradient of the oracle (oracle_ad_gradient): returns
    (J, dJ/dw0 [9][nz][ny][nx], dJ/d(mu_s, mu_d, Dc, C0, T0x, T0y, T0z)
    [nfault][7][nz][ny]). Small grids only (one thread, a tape of every
    operation)."""
    nz, ny, nx = shape_ext
    w0 = np.ascontiguousarray(w0, np.float64)
    sta = np.ascontiguousarray(sta_ijk, np.int32).reshape(-1, 3)
    res = np.ascontiguousarray(res, np.float64)
    assert res.shape == (sta.shape[0], nt, 9)
    nrec = 0
    rec = resf = None
    if rec_fjk is not None:
        rec = np.ascontiguousarray(rec_fjk, np.int32).reshape(-1, 3)
        resf = np.ascontiguousarray(res_f, np.float64)
        nrec = rec.shape[0]
        assert resf.shape == (nrec, nt, 9)
    gw = np.zeros((9, nz, ny, nx), np.float64)
    gp = np.zeros((max(nfault, 1), 7, nz, ny), np.float64)
    J = C.c_double()
    load().oracle_ad_gradient(C.byref(desc), int(bool(fix_q1)), _p(w0), int(nt), sta.shape[0], _p(sta), _p(res), nrec,
                              None if rec is None else _p(rec), None if resf is None else _p(resf), _p(gw), _p(gp),
                              C.byref(J))
    return J.value, gw, gp[:nfault]
