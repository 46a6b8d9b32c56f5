"""A small rupture differentiated by the FP64 reverse-mode oracle
(oracle.ad_gradient): the scenario, its stations and fault receivers, and
seeded residuals, shared by the CPU pin of that reference adjoint and the
GPU comparison."""
from __future__ import annotations

import numpy as np

NT = 14
PARAMS = ["mu_s", "mu_d", "Dc", "C0", "Tx", "Ty", "Tz"]


def scenario(imethod: int = 2, locked: bool = False):
    from cgfd import configs
    sc = configs.small(ni=20, nj=28, nk=18, pml=3, imethod=imethod, dc=0.05)
    if locked:
        sc.faults[0].fields["mu_s"][:] = 1000.0
    return sc


def records(sc):
    """Ghost-inclusive stations (i, j, k) and fault receivers (fault, j, k)."""
    i0 = sc.faults[0].x_index
    j1, j2, _, _ = sc.faults[0].grid
    jc = (j1 + j2) // 2
    sta = np.array([[i0 - 1, jc, sc.nk - 1], [i0 + 2, jc + 1, sc.nk - 2], [3, 5, sc.nk - 1], [i0, jc - 2, sc.nk - 3],
                    [i0 + 4, jc + 3, sc.nk - 1]], np.int32) + 3
    rec = np.array([[0, jc + 3 + dj, sc.nk - 1 - dk + 3] for dj in (-2, 0, 2) for dk in (1, 3, 5)], np.int32)
    return sta, rec


def residuals(sc, nsta, nrec, seed=3):
    rng = np.random.default_rng(seed)
    res = rng.standard_normal((nsta, NT, 9))
    res[:, :, 3:] *= 1e-6
    resf = np.zeros((nrec, NT, 9))
    resf[:, :, 3] = rng.standard_normal((nrec, NT))
    resf[:, :, 6] = rng.standard_normal((nrec, NT))
    resf[:, 0] = 0.0          # sample 0 precedes the recorded steps
    return res, resf


def initial_wavefield(sc, seed=4, amp=1e-3):
    rng = np.random.default_rng(seed)
    nz, ny, nx = sc.shape_ext
    w0 = np.zeros((9, nz, ny, nx))
    w0[:3, 3:-3, 3:-3, 3:-3] = amp * rng.standard_normal((3, sc.nk, sc.nj, sc.ni))
    return w0.astype(np.float32).astype(np.float64)


def reference(sc, w0, sta, res, rec, resf):
    import oracle
    from oracle.setup_ref import SetupRef
    ref = SetupRef(sc)
    return oracle.ad_gradient(ref.desc(), sc.shape_ext, len(sc.faults), w0, NT, sta, res, rec, resf)


def misfit_fp64(sc, w0, sta, res, rec, resf):
    """J of the FP64 oracle (the function the reference differentiates)."""
    import oracle
    o = oracle.Oracle.from_scenario(sc, double=True)
    o.write_wavefield(w0)
    J = 0.0
    for t in range(NT):
        w = o.read_wavefield()
        f = o.read_fault(0)
        for s, (i, j, k) in enumerate(sta):
            J += float(np.dot(res[s, t], w[:, k, j, i]))
        for r, (_, j, k) in enumerate(rec):
            J += float(np.dot(resf[r, t, 3:9], f[3:9, k, j]))
        o.run(1)
    o.close()
    return J
