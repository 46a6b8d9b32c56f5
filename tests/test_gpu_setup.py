"""The setup's device path (cgfd_setup_params.device_setup = 1): the metric
(gd_curv_metric_cal, gd_t.cu:209-304) and the fault coefficients
(fault_coef_cal, fault_info.cu:155-845) computed by the HIP kernels of
csrc/k_setup.hip. They run the host loops' per-point arithmetic with FMA
contraction off, IEEE division and square root and FP64 cofactor inverses,
so every array must equal the host setup's bit for bit (the host setup is
itself pinned to closed forms and to an independent FP64 restatement in
test_setup_ref.py); a Cartesian grid is also checked against the closed
forms directly.
"""
import numpy as np
import pytest

from helpers import progress

pytestmark = pytest.mark.gpu


def _cases():
    from cgfd import configs
    return {
        "small curvilinear method 1": configs.small(imethod=1, curvilinear=True),
        "small curvilinear method 2": configs.small(imethod=2, curvilinear=True),
        "config2 at 1/10 scale": configs.config2(scale=0.1, steps=10),
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_device_setup_equals_host_setup(built_libs, name):
    from cgfd import model
    from test_input_readers import _setup_arrays
    sc = _cases()[name]
    progress(f"[device setup] {name}: {sc.ni}x{sc.nj}x{sc.nk}")
    st_h = model.Setup(sc)
    st_d = model.Setup(sc, device_setup=True)
    a, b = _setup_arrays(st_h, sc), _setup_arrays(st_d, sc)
    st_h.close()
    st_d.close()
    assert a.keys() == b.keys()
    bad = []
    for k in a:
        if not np.array_equal(a[k], b[k]):
            x, y = np.asarray(a[k], np.float64), np.asarray(b[k], np.float64)
            bad.append(f"{k}: {int(np.sum(x != y))} of {x.size} differ, max rel "
                       f"{np.abs(x - y).max() / max(np.abs(x).max(), 1e-30):.2e}")
    assert not bad, "; ".join(bad[:8])


def test_device_setup_cartesian_closed_forms(built_libs):
    """Cartesian grid of spacing h, isotropic medium: jac = h^3, grad xi =
    (1/h, 0, 0) etc., D11 = h diag(l+2m, m, m), n = x, s1 = y, s2 = z."""
    from cgfd import configs, model
    from test_input_readers import _setup_arrays
    sc = configs.small(imethod=2, curvilinear=False)
    h = float(sc.y[0, 1, 0] - sc.y[0, 0, 0])
    st = model.Setup(sc, device_setup=True)
    got = _setup_arrays(st, sc)
    st.close()
    met = got["metric"][:, 3:-3, 3:-3, 3:-3]
    assert np.allclose(met[0], h ** 3, rtol=2e-6)
    eye = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    for c in range(9):
        assert np.allclose(met[1 + c], eye[c] / h, rtol=2e-6, atol=1e-9), c
    nz, ny, nx = sc.shape_ext
    i0 = sc.faults[0].x_index
    lam = float(sc.lam[5, 5, i0])
    mu = float(sc.mu[5, 5, i0])
    pf = (5 + 3) + (5 + 3) * ny
    # D stores D21..D33 of side 1 then side 2; D11 itself is not stored, but
    # D22 (index 1) = h diag(m, l+2m, m) on a Cartesian grid
    d22 = got["0.D1"][pf * 9:(pf + 1) * 9].reshape(3, 3)
    ref = h * np.diag([mu, lam + 2 * mu, mu])
    assert np.allclose(d22, ref, rtol=2e-6, atol=1e-9 * np.abs(ref).max())
    vs1 = got["0.vec_s1"][pf * 3:(pf + 1) * 3]
    vs2 = got["0.vec_s2"][pf * 3:(pf + 1) * 3]
    assert np.allclose(vs1, [0, 1, 0], atol=1e-6) and np.allclose(vs2, [0, 0, 1], atol=1e-6)
