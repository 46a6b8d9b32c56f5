"""The GPU discrete adjoint against the FP64 reference adjoint (reverse-mode
differentiation of the FP64 oracle, pinned against FP64 central differences
in test_oracle_adjoint.py) on a small rupture with PML, a free surface,
stations and fault receivers: whole gradient fields with respect to the
initial wavefield and to every friction and stress parameter of every fault
node, both dynamic methods, on a rupturing fault and on a locked one (the
purely linear path: wave2fault, traction columns, trial, fault2wave, split
nodes, hT1 recurrence, free-surface rows)."""
import numpy as np
import pytest

import adjoint_ref as A
from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("imethod,locked", [(2, False), (1, False), (2, True), (1, True)])
def test_gpu_adjoint_matches_fp64_reference_adjoint(built_libs, imethod, locked):
    from cgfd import model
    sc = A.scenario(imethod, locked)
    sta, rec = A.records(sc)
    res, resf = A.residuals(sc, len(sta), len(rec))
    res = res.astype(np.float32)
    resf = resf.astype(np.float32)
    w0 = A.initial_wavefield(sc)
    st = model.Setup(sc)
    core = model.Core(st)
    core.set_stations(sta, A.NT)
    core.set_fault_receivers(rec, A.NT)
    core.adjoint_forward(A.NT)
    core.write_wavefield(w0.astype(np.float32))
    core.run(A.NT)
    g, gf = core.adjoint_fault(res, resf)
    core.close()
    st.close()
    J, gw, gp = A.reference(sc, w0, sta, res.astype(np.float64), rec, resf.astype(np.float64))
    phys = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    out = {"w0": rel_l2(g[phys], gw[phys])}
    for q, name in enumerate(A.PARAMS):
        ref = gp[0, q, 3:-3, 3:-3]
        if np.abs(ref).max() > 0:
            out[name] = rel_l2(gf[0, q], ref)
    print(f"[adjoint vs FP64 reference, method {imethod}, locked {locked}] J {J:.6e}; field rel L2: "
          + ", ".join(f"{k} {v:.1e}" for k, v in out.items()), flush=True)
    assert len(out) >= (1 if locked else 7)   # a locked fault: only the initial-state gradient
    tol = 1e-4 if locked else 1e-3
    assert max(out.values()) <= tol, out
