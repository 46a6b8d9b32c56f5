"""Comparison helpers shared by the parity tests."""
import numpy as np


def interior(w, g=3):
    """Physical points of an extended array (last three axes)."""
    return w[..., g:-g, g:-g, g:-g]


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def fault_phys(f, g=3):
    """Physical nodes of fault-plane arrays (last two axes nz, ny)."""
    return f[..., g:-g, g:-g]


def progress(msg):
    """A progress line: printed, and where GRAFT_REPO_ROOT is set also
    appended to that run's progress log, so a long test shows it is alive
    even when pytest captures its output."""
    import os
    print(msg, flush=True)
    root = os.environ.get("GRAFT_REPO_ROOT")
    if root:
        try:
            d = os.path.join(root, "gpurun_out")
            os.makedirs(d, exist_ok=True)
            with open(os.path.join(d, "test_progress.log"), "a") as f:
                f.write(msg + "\n")
        except OSError:
            pass
