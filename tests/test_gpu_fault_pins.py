"""The closed-form fault cases of test_oracle_fault_pins.py run on the GPU
core through the C ABI: uniform stress drop against the 1-D solution and
the FP32 oracle, uniform strain rate exact after one step (split-node
stress rows, both methods, free surface, sheared grid), and the Txz/Tyz
swap in the fault-plane column with and without the correction."""
import numpy as np
import pytest

import fault_pins as fp
import pins
from helpers import fault_phys, interior, rel_l2
from test_oracle_fault_pins import STRESS_DROP_STEPS, stress_drop_scenario

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["strike", "dip"])
@pytest.mark.parametrize("imethod", [1, 2])
def test_gpu_uniform_stress_drop(built_libs, mode, imethod):
    from cgfd import model
    import oracle
    sc = stress_drop_scenario(mode, imethod)
    print(f"stress drop {mode} method {imethod} {sc.ni}x{sc.nj}x{sc.nk}", flush=True)
    st = model.Setup(sc)
    core = model.Core(st)
    core.run(STRESS_DROP_STEPS)
    fo, fs, w = core.read_fault(0), core.read_fault_state(0), core.read_wavefield()
    r = fp.check_stress_drop(fo, fs, w, sc, STRESS_DROP_STEPS)
    assert r["ok"], r
    orc = oracle.Oracle.from_scenario(sc, double=False)
    orc.run(STRESS_DROP_STEPS)
    assert rel_l2(interior(w), interior(orc.read_wavefield())) <= 1e-3
    assert rel_l2(fault_phys(fo[6]), fault_phys(orc.read_fault(0)[6])) <= 1e-3
    core.close()
    orc.close()
    st.close()


# FP32 arithmetic on a linear field: differences of velocities ~1 m/s
# resolve the gradient to ~1e-5 relative
SR_TOL = {True: 2e-4, False: 1e-3}


def _gpu_strain_rate(imethod, cart, free_top, compat=1):
    from cgfd import model
    A = np.diag([100.0, 100.0, 100.0]) if cart else pins.affine_matrix(100.0, seed=4)
    sc = fp.strain_rate_scenario(imethod, A, free_top=free_top)
    G = fp.strain_rate_gradient(A, free_top)
    st = model.Setup(sc)
    core = model.Core(st, compat_flags=compat)
    w0, f0, t1, w1, f1 = fp.strain_rate_state(sc, A, G)
    core.write_wavefield(w0)
    core.write_fault_state(f0, 0, t1)
    core.run(1)
    out = (core.read_wavefield(), core.read_fault_state(0), core.read_fault(0))
    core.close()
    st.close()
    return sc, out, (w1, f1)


@pytest.mark.parametrize("cart", [True, False])
@pytest.mark.parametrize("free_top", [True, False])
@pytest.mark.parametrize("imethod", [1, 2])
def test_gpu_uniform_strain_rate_is_exact(built_libs, cart, free_top, imethod):
    print(f"strain rate cart={cart} free={free_top} method {imethod}", flush=True)
    sc, (w, f, fo), (w1, f1) = _gpu_strain_rate(imethod, cart, free_top)
    top = 5 if (free_top and imethod == 2) else 0
    e = fp.strain_rate_errors(w, f, w1, f1, sc, top_skip=top)
    assert max(e) <= SR_TOL[cart], e
    assert np.abs(fo[3]).max() < 1e-6


@pytest.mark.parametrize("fix", [0, 1])
def test_gpu_plane_column_txz_tyz_swap(built_libs, fix):
    sc, (w, f, _), (w1, f1) = _gpu_strain_rate(1, False, False, compat=fix)
    i0 = sc.faults[0].x_index + 3
    c = (slice(3 + 14, 3 + sc.nk - 14), slice(3 + 14, 3 + sc.nj - 14), i0)
    nrm = np.abs(w1[3:]).max()
    exp_tyz, exp_txz = (w1[6][c], w1[7][c]) if fix else (w1[7][c], w1[6][c])
    assert np.abs(w[6][c] - exp_tyz).max() <= 1e-3 * nrm
    assert np.abs(w[7][c] - exp_txz).max() <= 1e-3 * nrm
    assert np.abs(w1[6][c] - w1[7][c]).max() > 0.1 * nrm
