"""TEST INFRASTRUCTURE ONLY -- an independent FP64 restatement of the host
setup that feeds the time loop.

numpy, vectorised over grid points and fault nodes, written from the
reference's setup routines and the mathematics they implement, not from the
product's C++ (gpu-cgfd3d-dynamic_amd/host/cgfd_setup.cpp). Tests use it two
ways: to check every array of the product's descriptor against a second
coding, and to build the oracle's descriptor without the product's setup, so
a mistake in one setup cannot reach both sides of a parity test.

    coordinates + ghosts   geometric_symmetry        gd_t.cu:665-759
    metric                 gd_curv_metric_cal        gd_t.cu:209-304
    media ghosts, 1/rho    mirror_symmetry           gd_t.cu:578-655
                           md_rho_to_slow            md_t.cu:568-582
    CFS-PML profiles       bdry_pml_set              bdry_t.cu:79-251, 291-324
                           bdry_cal_abl_len_dh       bdry_t.cu:329-445
    fault coefficients     fault_coef_cal            fault_info.cu:97-850
    fault masks, T0 frame  fault_set                 fault_info.cu:967-1074

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

G = 3                                   # ghost layers per side
# 6th-order centred first derivative (fd_t.h:126-140); d_4 = 0
D6 = np.array([-0.01666667, 0.15, -0.75, 0.0, 0.75, -0.15, 0.01666667])
CONSPD, CONSPB, CONSPA = 2.0, 2.0, 1.0   # bdry_t.cu:15-17


# ---------------------------------------------------------------- ghosts
def _extend(v: np.ndarray, odd: bool) -> np.ndarray:
    """(nk, nj, ni) -> ghost-padded (nk+6, nj+6, ni+6): per axis in the
    order x, y, z, ghost g = 2 v[b] - v[2b - g] (odd, linear extrapolation)
    or v[2b - g] (even mirror), b the boundary node."""
    out = np.zeros(tuple(n + 2 * G for n in v.shape), np.float64)
    out[G:-G, G:-G, G:-G] = v
    for ax in (2, 1, 0):
        n = out.shape[ax]
        b1, b2 = G, n - G - 1
        for g in range(G):
            src1 = np.take(out, 2 * b1 - g, axis=ax)
            src2 = np.take(out, 2 * b2 - (n - 1 - g), axis=ax)
            if odd:
                src1 = 2 * np.take(out, b1, axis=ax) - src1
                src2 = 2 * np.take(out, b2, axis=ax) - src2
            idx1 = [slice(None)] * 3
            idx1[ax] = g
            idx2 = [slice(None)] * 3
            idx2[ax] = n - 1 - g
            out[tuple(idx1)] = src1
            out[tuple(idx2)] = src2
    return out


def _dcenter(a: np.ndarray, ax: int) -> np.ndarray:
    """6th-order centred derivative at the physical points of a padded array
    along axis ax (0 = k, 1 = j, 2 = i)."""
    phys = [slice(G, -G)] * 3
    acc = np.zeros(tuple(s - 2 * G for s in a.shape))
    for m, c in zip(range(-3, 4), D6):
        if c == 0.0:
            continue
        sl = list(phys)
        sl[ax] = slice(G + m, a.shape[ax] - G + m)
        acc += c * a[tuple(sl)]
    return acc


# ---------------------------------------------------------------- metric
def metric(x: np.ndarray, y: np.ndarray, z: np.ndarray):
    """Padded coordinates and the 10 metric arrays (jac, xi_x..zeta_z),
    ghost-padded, from physical coordinates (nk, nj, ni)."""
    X = [_extend(np.asarray(c, np.float64), True) for c in (x, y, z)]
    # dr/dxi, dr/deta, dr/dzeta: derivatives along i, j, k
    r_xi = np.stack([_dcenter(c, 2) for c in X], -1)
    r_et = np.stack([_dcenter(c, 1) for c in X], -1)
    r_zt = np.stack([_dcenter(c, 0) for c in X], -1)
    jac = np.einsum("...a,...a->...", np.cross(r_xi, r_et), r_zt)
    g_xi = np.cross(r_et, r_zt) / jac[..., None]
    g_et = np.cross(r_zt, r_xi) / jac[..., None]
    g_zt = np.cross(r_xi, r_et) / jac[..., None]
    phys = [jac] + [g[..., a] for g in (g_xi, g_et, g_zt) for a in range(3)]
    met = np.stack([_extend(p, True) for p in phys])
    return np.stack(X), met


# ---------------------------------------------------------------- PML
def pml(sc, coords: np.ndarray):
    """Per face [dim][side]: on, ghost-inclusive index ranges and the
    profiles A, B, D (after D/B, A + D, 1/B), nlay + 1 values each."""
    ni1, nj1, nk1 = G, G, G
    ni2, nj2, nk2 = G + sc.ni - 1, G + sc.nj - 1, G + sc.nk - 1
    faces = {}
    for dim in range(3):
        for side in range(2):
            nl = int(sc.pml_layers[dim][side])
            on = nl > 0 and not (dim == 2 and side == 1 and sc.free_top)
            if not on:
                nl = 0
            r = [ni1, ni2, nj1, nj2, nk1, nk2]
            lo, hi = 2 * dim, 2 * dim + 1
            if side == 0:
                r[hi] = r[lo] + nl
            else:
                r[lo] = r[hi] - nl
            face = {"on": on, "range": r}
            if on:
                # average spacing of consecutive nodes along the face's axis
                sl = (slice(r[4], r[5] + 1), slice(r[2], r[3] + 1), slice(r[0], r[1] + 1))
                c = coords[(slice(None),) + sl]
                ax = {0: 3, 1: 2, 2: 1}[dim]
                seg = np.sqrt((np.diff(c, axis=ax) ** 2).sum(axis=0))
                dh = float(np.float32(seg.sum() / seg.size))
                L0 = float(np.float32(dh * nl))
                nlay = float(nl)
                R = 10.0 ** (-((np.log10(nlay) - 1.0) / np.log10(2.0) + 4.0))
                dmax = -float(sc.pml_velocity) / (2.0 * L0) * np.log(R) * (CONSPD + 1.0)
                n = np.arange(nl + 1)
                Lx = n * dh
                d = dmax * (Lx / L0) ** CONSPD
                a = float(sc.pml_alpha_max) * (1.0 - (Lx / L0) ** CONSPA)
                b = 1.0 + (float(sc.pml_beta_max) - 1.0) * (Lx / L0) ** CONSPB
                Dv = d / b
                Av = a + Dv
                Bv = 1.0 / b
                order = n[::-1] if side == 0 else n        # grid index from the low side
                A_, B_, D_ = np.empty(nl + 1), np.empty(nl + 1), np.empty(nl + 1)
                A_[order], B_[order], D_[order] = Av, Bv, Dv
                face.update(A=A_, B=B_, D=D_)
            faces[(dim, side)] = face
    return faces


# ---------------------------------------------------------------- fault coefficients
def dmat(ea: np.ndarray, eb: np.ndarray, lam: np.ndarray, mu: np.ndarray, jac: np.ndarray) -> np.ndarray:
    """Traction-rate matrix of the plane with normal e_a for velocity
    derivatives along b: J (lam e_a e_b^T + mu e_b e_a^T + mu (e_a.e_b) I),
    [..., 3, 3] (the 18 blocks of fault_info.cu:190-369)."""
    outer_ab = ea[..., :, None] * eb[..., None, :]
    outer_ba = eb[..., :, None] * ea[..., None, :]
    dot = np.einsum("...a,...a->...", ea, eb)
    M = lam[..., None, None] * outer_ab + mu[..., None, None] * (outer_ba + dot[..., None, None] * np.eye(3))
    return jac[..., None, None] * M


def fault_coefs(met: np.ndarray, lam: np.ndarray, mu: np.ndarray, i0: int, k_top: int, plant: int = 0):
    """Coefficients of fault_coef_cal on the plane i = i0 (ghost-inclusive)
    for every (k, j) of the padded grid, both sides sharing one medium.
    plant = 1 exchanges lambda and mu in the off-diagonal (shear) terms of
    every traction matrix -- a deliberate mistake for the test that shows the
    closed-form checks catch it."""
    jac = met[0][:, :, i0]
    e = [np.stack([met[1 + 3 * a + c][:, :, i0] for c in range(3)], -1) for a in range(3)]
    la, mu_ = lam[:, :, i0], mu[:, :, i0]
    if plant == 1:
        def D(a, b):
            ea, eb = e[a], e[b]
            M = dmat(ea, eb, la, mu_, jac)
            off = ~np.eye(3, dtype=bool)
            swapped = dmat(ea, eb, mu_, la, jac)
            M[..., off] = swapped[..., off]
            return M
    else:
        def D(a, b):
            return dmat(e[a], e[b], la, mu_, jac)
    D11, D12, D13 = D(0, 0), D(0, 1), D(0, 2)
    Dst = [D(a, b) for a in (1, 2) for b in range(3)]          # D21 D22 D23 D31 D32 D33
    iD11 = np.linalg.inv(D11)
    out = {}
    out["D"] = Dst + Dst                                      # side 1 (minus) then side 2 (plus)
    # method 1 (fault_info.cu:404-526): the same five products on both sides
    m1 = [iD11 @ D11, iD11 @ D12, iD11 @ D13, iD11 @ D12, iD11 @ D13]
    out["P2M"] = m1
    out["M2P"] = m1
    # method 2 (fault_info.cu:528-607)
    out["M2"] = [iD11, iD11 @ D12, iD11 @ D13, iD11, iD11 @ D12, iD11 @ D13]
    # frame (fault_info.cu:609-654): strike = dr/deta normalised, normal =
    # grad xi normalised, s2 = n x s1
    M3 = np.stack([e[0], e[1], e[2]], -2)                     # rows grad xi, grad eta, grad zeta
    r_et = np.linalg.inv(M3)[..., :, 1]
    s1 = r_et / np.linalg.norm(r_et, axis=-1, keepdims=True)
    n = e[0] / np.linalg.norm(e[0], axis=-1, keepdims=True)
    s2 = np.cross(n, s1)
    out["vec_n"], out["vec_s1"], out["vec_s2"] = n, s1, s2
    # free-surface row at k_top (fault_info.cu:656-845)
    k = k_top
    D31, D32, D33 = Dst[3][k], Dst[4][k], Dst[5][k]
    iD33 = np.linalg.inv(D33)
    Vx2Vz = iD33 @ -D31
    Vy2Vz = iD33 @ -D32
    vn = e[0][k]                                               # not normalised here
    s2b = np.cross(vn, r_et[k])
    s2b /= np.linalg.norm(s2b, axis=-1, keepdims=True)
    nv = np.linalg.norm(vn, axis=-1)
    h = [np.einsum("...a,...a->...", s2b, e[a][k]) / nv for a in range(3)]
    To = s2b[..., :, None] * vn[..., None, :]
    mk = mu_[k]
    D1f = [D11[k] + (mk * h[0])[..., None, None] * To, D12[k] + (mk * h[1])[..., None, None] * To,
           D13[k] + (mk * h[2])[..., None, None] * To]
    Vxf = D1f[0] + D1f[2] @ Vx2Vz                              # both sides alike
    Vyf = D1f[1] + D1f[2] @ Vy2Vz
    iVxf = np.linalg.inv(Vxf)
    P2M1f, P2M2f, P2M3f = iVxf @ Vxf, iVxf @ Vyf, iVxf @ Vyf
    out["F"] = [Vx2Vz, Vy2Vz, Vx2Vz, Vy2Vz, P2M1f, P2M2f, P2M3f, P2M1f, P2M2f, P2M3f, iVxf, P2M3f, iVxf, P2M3f]
    return out


def fault_set(fields: dict, grid, coefs, nj: int, nk: int, margin: int, free_top: bool):
    """Masks and the initial traction frame of fault_set (0-based fault_grid
    j1, j2, k1, k2 of physical nodes); arrays over the padded plane (nz, ny)."""
    ny, nz = nj + 2 * G, nk + 2 * G
    gj = np.arange(ny)[None, :] - G
    gk = np.arange(nz)[:, None] - G
    j1, j2, k1, k2 = grid
    rupy = np.broadcast_to(~((gj <= j1 + 3) | (gj >= j2 - 3)), (nz, ny)).astype(np.int32)
    rupz = np.broadcast_to(~((gk <= k1 + 3) | (gk >= k2 - 3)), (nz, ny)).astype(np.int32)
    inside = (gj >= margin) & (gj < nj - margin) & (gk >= margin)
    if not free_top:
        inside = inside & (gk < nk - margin)
    united = np.where(inside, 0, 1).astype(np.int32)
    fgrid = ((gj >= j1 + 3) & (gj <= j2 - 3) & (gk >= k1 + 3) & (gk <= k2 - 3)).astype(np.int32)
    T0 = np.zeros((nz, ny, 3))
    for a, nm in enumerate(("Tx", "Ty", "Tz")):
        T0[G:-G, G:-G, a] = fields[nm]
    out = {"united": united, "faultgrid": fgrid, "rupy": rupy, "rupz": np.ascontiguousarray(rupz)}
    out["Tn0"] = np.einsum("...a,...a->...", T0, coefs["vec_n"])
    out["Ts10"] = np.einsum("...a,...a->...", T0, coefs["vec_s1"])
    out["Ts20"] = np.einsum("...a,...a->...", T0, coefs["vec_s2"])
    return out


# ---------------------------------------------------------------- everything
class SetupRef:
    """The whole setup of a cgfd.model.Scenario in FP64, and a boundary
    descriptor (cgfd.lib.Desc) over FP32 copies of it for the oracle."""

    def __init__(self, sc, plant: int = 0):
        self.sc = sc
        self.coords, self.met = metric(sc.x, sc.y, sc.z)
        self.lam = _extend(np.asarray(sc.lam, np.float64), False)
        self.mu = _extend(np.asarray(sc.mu, np.float64), False)
        rho = _extend(np.asarray(sc.rho, np.float64), False)
        self.rho = rho
        self.slw = np.where(rho > 1e-10, 1.0 / np.where(rho > 1e-10, rho, 1.0), 0.0)
        self.pml = pml(sc, self.coords)
        k_top = G + sc.nk - 1
        self.faults = []
        for f in sc.faults:
            i0 = f.x_index + G
            co = fault_coefs(self.met, self.lam, self.mu, i0, k_top, plant=plant)
            fs = fault_set(f.fields, f.grid, co, sc.nj, sc.nk, sc.united_margin, sc.free_top)
            co.update(fs)
            co["rho_f"] = np.stack([rho[:, :, i0], rho[:, :, i0]])
            self.faults.append(co)
        self._keep = []

    # -------------------------------------------------------- descriptor
    def _f32(self, a, dtype=np.float32):
        a = np.ascontiguousarray(a, dtype)
        self._keep.append(a)
        return a.ctypes.data_as(C.c_void_p)

    def desc(self, compat_flags: int = 0):
        """A cgfd.lib.Desc over FP32 copies of the arrays (kept alive by this
        object) in the host layouts the C ABI documents."""
        from cgfd import lib as L
        sc = self.sc
        d = L.Desc()
        d.ni, d.nj, d.nk = sc.ni, sc.nj, sc.nk
        d.metric = self._f32(self.met)
        d.lambda_ = self._f32(self.lam)
        d.mu = self._f32(self.mu)
        d.rho = self._f32(self.slw)
        d.free_top = int(sc.free_top)
        for (dim, side), f in self.pml.items():
            d.pml_is_sides[dim][side] = int(f["on"])
            r = f["range"]
            d.pml_ni1[dim][side], d.pml_ni2[dim][side] = r[0], r[1]
            d.pml_nj1[dim][side], d.pml_nj2[dim][side] = r[2], r[3]
            d.pml_nk1[dim][side], d.pml_nk2[dim][side] = r[4], r[5]
            if f["on"]:
                d.pml_A[dim][side] = self._f32(f["A"])
                d.pml_B[dim][side] = self._f32(f["B"])
                d.pml_D[dim][side] = self._f32(f["D"])
        d.imethod = sc.imethod
        d.number_fault = len(sc.faults)
        for fid, (spec, co) in enumerate(zip(sc.faults, self.faults)):
            fd = d.fault[fid]
            fd.fault_x_index = spec.x_index
            names = ("T0x", "T0y", "T0z", "mu_s", "mu_d", "Dc", "C0")
            for nm, key in zip(names, ("Tx", "Ty", "Tz", "mu_s", "mu_d", "Dc", "C0")):
                pad = np.zeros((sc.nk + 2 * G, sc.nj + 2 * G))
                pad[G:-G, G:-G] = spec.fields[key]
                setattr(fd, nm, self._f32(pad))
            fd.united = self._f32(co["united"], np.int32)
            fd.faultgrid = self._f32(co["faultgrid"], np.int32)
            fd.rup_index_y = self._f32(co["rupy"], np.int32)
            fd.rup_index_z = self._f32(co["rupz"], np.int32)
            fd.Tn0, fd.Ts10, fd.Ts20 = self._f32(co["Tn0"]), self._f32(co["Ts10"]), self._f32(co["Ts20"])
            fd.rho_f = self._f32(co["rho_f"])
            fd.vec_s1 = self._f32(co["vec_s1"])
            fd.vec_s2 = self._f32(co["vec_s2"])
            for q in range(12):
                fd.D[q] = self._f32(co["D"][q])
            for q in range(5):
                fd.matPlus2Min[q] = self._f32(co["P2M"][q])
                fd.matMin2Plus[q] = self._f32(co["M2P"][q])
            for q in range(6):
                fd.matM2[q] = self._f32(co["M2"][q])
            for q in range(14):
                fd.matF[q] = self._f32(co["F"][q])
        d.dt, d.t0, d.nt_total = sc.dt, sc.t0, sc.nt_total
        d.device_id = 0
        d.compat_flags = compat_flags
        return d
