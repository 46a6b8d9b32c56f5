"""Shared fixtures. GPU tests are marked ``gpu``; everything else runs on CPU."""
import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "gpu-cgfd3d-dynamic_amd"))
sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def built_libs():
    from cgfd import build
    # rebuild when a source is newer than the library, so a test never runs
    # a binary older than the tree it ships with
    if build.core_is_stale():
        build.build_core()
    if build.oracle_is_stale():
        build.build_oracle()
    return build.LIB_PATH, build.ORACLE_LIB


def pytest_runtest_logstart(nodeid, location):
    """Where GRAFT_REPO_ROOT names the repository root of a remote run,
    append each test's start to that run's progress log (a run shows it is
    alive even while pytest captures the output)."""
    root = os.environ.get("GRAFT_REPO_ROOT")
    if not root:
        return
    try:
        d = os.path.join(root, "gpurun_out")
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "test_progress.log"), "a") as f:
            f.write(f"start {nodeid}\n")
    except OSError:
        pass
