"""The FP64 reference adjoint (reverse-mode differentiation of the FP64
oracle, oracle_ad_gradient) against FP64 central differences of the same
oracle on a small rupture with PML, a free surface, one slipping fault,
stations and fault receivers: the gradient with respect to the initial
wavefield and to mu_s, mu_d, Dc and the shear traction along seeded
directions. The steps are a few FP32 units in the last place of each
parameter (the quotient uses the perturbation actually taken), small enough
that no node's rupture onset or Dc crossing moves by a time step inside the
difference interval."""
import copy

import numpy as np

import adjoint_ref as A


def test_reference_adjoint_matches_fp64_central_differences(built_libs):
    sc = A.scenario()
    sta, rec = A.records(sc)
    res, resf = A.residuals(sc, len(sta), len(rec))
    w0 = A.initial_wavefield(sc)
    J, gw, gp = A.reference(sc, w0, sta, res, rec, resf)
    J2 = A.misfit_fp64(sc, w0, sta, res, rec, resf)
    assert abs(J - J2) <= 1e-12 * abs(J2)
    rng = np.random.default_rng(11)
    j1, j2, k1, k2 = sc.faults[0].grid
    errs = {}
    for name, rel in (("mu_s", 2e-6), ("mu_d", 2e-6), ("Dc", 2e-6), ("Ty", 2e-6)):
        d = np.zeros((sc.nk, sc.nj))
        d[k1:k2 + 1, j1:j2 + 1] = rng.standard_normal((k2 - k1 + 1, j2 - j1 + 1))
        base = sc.faults[0].fields[name].astype(np.float64)
        eps = rel * np.abs(base[k1:k2 + 1, j1:j2 + 1]).max()
        sp, sm = copy.deepcopy(sc), copy.deepcopy(sc)
        sp.faults[0].fields[name] = (base + eps * d).astype(np.float32)
        sm.faults[0].fields[name] = (base - eps * d).astype(np.float32)
        step = sp.faults[0].fields[name].astype(np.float64) - sm.faults[0].fields[name].astype(np.float64)
        fd = A.misfit_fp64(sp, w0, sta, res, rec, resf) - A.misfit_fp64(sm, w0, sta, res, rec, resf)
        ad = float(np.sum(gp[0, A.PARAMS.index(name), 3:-3, 3:-3] * step))
        errs[name] = abs(ad - fd) / abs(fd)
    dw = np.zeros_like(w0)
    dw[:, 3:-3, 3:-3, 3:-3] = rng.standard_normal((9, sc.nk, sc.nj, sc.ni)) * np.array([1e-7] * 3 + [1e2] * 6)[:, None, None, None]
    fd = A.misfit_fp64(sc, w0 + dw, sta, res, rec, resf) - A.misfit_fp64(sc, w0 - dw, sta, res, rec, resf)
    errs["w0"] = abs(float(np.sum(gw * 2 * dw)) - fd) / abs(fd)
    print("reference adjoint vs FP64 central differences:", {k: f"{v:.1e}" for k, v in errs.items()})
    assert max(errs.values()) <= 1e-6, errs
    # the slipping fault is exercised
    assert np.abs(gp[0, A.PARAMS.index("mu_d")]).max() > 0.0
