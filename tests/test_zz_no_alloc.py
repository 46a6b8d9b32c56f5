"""cgfd_core_read_snapshot_noalloc (no device allocation: a fixed 32 MiB
device array per GPU) against cgfd_core_read_snapshot on two calls in a
row, the second with a smaller box, and its size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(core, box, flags):
    from cgfd import lib as L
    b = np.ascontiguousarray(box, np.int32)
    nf = (3 if flags & 1 else 0) + (6 if flags & 2 else 0) + (6 if flags & 4 else 0)
    out = np.full((nf, int(b[7]), int(b[4]), int(b[1])), np.nan, np.float32)
    rc = core.lib.cgfd_core_read_snapshot_noalloc(core.handle, L.ptr(b), flags, L.ptr(out))
    return rc, out


def test_noalloc_snapshot_matches_allocating_reader(built_libs):
    from cgfd import configs, model
    sc = configs.small(imethod=2, curvilinear=True)
    st = model.Setup(sc)
    core = model.Core(st)
    core.run(20)
    big = (3, sc.ni, 1, 3, sc.nj, 1, 3, sc.nk, 1)                 # the whole grid, 15 fields
    small = (5, 11, 2, 7, 9, 3, 10, 6, 2)                         # a strided sub-box
    for box, flags in ((big, 7), (small, 3)):
        print(f"[noalloc] box {box} flags {flags}", flush=True)
        ref = core.read_snapshot(box, flags)
        rc, out = _noalloc(core, box, flags)
        assert rc == 0
        assert np.array_equal(out, ref)
    core.run(3)
    rc, out = _noalloc(core, small, 7)
    assert rc == 0 and np.array_equal(out, core.read_snapshot(small, 7))
    core.close()
    st.close()


def test_noalloc_snapshot_rejects_boxes_beyond_its_maximum(built_libs):
    from cgfd import configs, lib as L, model
    sc = configs.small(ni=100, nj=100, nk=60, pml=5, imethod=2)   # 600k nodes x 15 fields = 9e6 > 8 Mi values
    st = model.Setup(sc)
    core = model.Core(st)
    box = np.array([3, sc.ni, 1, 3, sc.nj, 1, 3, sc.nk, 1], np.int32)
    out = np.zeros(1, np.float32)
    rc = core.lib.cgfd_core_read_snapshot_noalloc(core.handle, L.ptr(box), 7, L.ptr(out))
    assert rc == -1                                                  # CGFD_ERR_ARG
    assert b"allocation-free maximum" in core.lib.cgfd_last_error()
    rc, ok = _noalloc(core, (3, sc.ni, 1, 3, sc.nj, 1, 3, sc.nk, 1), 3)   # 3.6e6 values fit
    assert rc == 0 and np.array_equal(ok, core.read_snapshot((3, sc.ni, 1, 3, sc.nj, 1, 3, sc.nk, 1), 3))
    core.close()
    st.close()
