"""Seeded synthetic configurations.

Axis convention (the reference only supports faults on constant-x planes,
par_t.cu:316-331): x = fault normal, y = strike, z = up, k = nk-1 is the
free surface. A configuration quoted as "X x Y x Z with the fault on the mid
y-plane" is laid out here with ni = Y (fault normal), nj = X (strike).

Random numbers come from numpy's PCG64 generator seeded per configuration;
the same arrays feed the GPU core and the CPU oracle.
"""
from __future__ import annotations

import numpy as np

from .model import FaultSpec, Scenario

MPA = 1.0e6


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def cartesian_coords(ni: int, nj: int, nk: int, dh: float, i0: int):
    """Flat-top Cartesian grid; x = 0 on the fault plane, z = 0 on the surface."""
    i = np.arange(ni, dtype=np.float64)
    j = np.arange(nj, dtype=np.float64)
    k = np.arange(nk, dtype=np.float64)
    x = np.broadcast_to(((i - i0) * dh)[None, None, :], (nk, nj, ni))
    y = np.broadcast_to(((j - nj // 2) * dh)[None, :, None], (nk, nj, ni))
    z = np.broadcast_to(((k - (nk - 1)) * dh)[:, None, None], (nk, nj, ni))
    return (np.ascontiguousarray(x, np.float32), np.ascontiguousarray(y, np.float32),
            np.ascontiguousarray(z, np.float32))


def hills(xs: np.ndarray, ys: np.ndarray, rng: np.random.Generator, n: int, hmax: float,
          wmin: float, wmax: float) -> np.ndarray:
    """Sum of n Gaussian hills with U(0,hmax) heights and U(wmin,wmax) widths."""
    topo = np.zeros(np.broadcast(xs, ys).shape)
    x0, x1, y0, y1 = xs.min(), xs.max(), ys.min(), ys.max()
    for _ in range(n):
        h = rng.uniform(0.0, hmax)
        w = rng.uniform(wmin, wmax)
        xc = rng.uniform(x0, x1)
        yc = rng.uniform(y0, y1)
        topo += h * np.exp(-((xs - xc) ** 2 + (ys - yc) ** 2) / (2.0 * w * w))
    return topo


def curvilinear_coords(ni: int, nj: int, nk: int, dh: float, i0: int, topo: np.ndarray,
                       dz_top: float, dz_bot: float):
    """Vertical columns (so every constant-i plane stays planar), topography
    on top, vertical spacing growing linearly from dz_top to dz_bot."""
    nint = nk - 1
    sp = dz_top + (dz_bot - dz_top) * np.arange(nint) / max(nint - 1, 1)   # from the surface down
    depth = np.concatenate([[0.0], np.cumsum(sp)])                         # depth of k-level from top
    zflat = -depth[::-1]                                                   # k = 0 bottom ... nk-1 top
    zb = zflat[0]
    i = np.arange(ni, dtype=np.float64)
    j = np.arange(nj, dtype=np.float64)
    x = np.broadcast_to(((i - i0) * dh)[None, None, :], (nk, nj, ni))
    y = np.broadcast_to(((j - nj // 2) * dh)[None, :, None], (nk, nj, ni))
    w = (zflat - zb) / (0.0 - zb)                                           # 0 at the bottom, 1 on top
    z = zflat[:, None, None] + topo[None, :, :] * w[:, None, None]
    return (np.ascontiguousarray(x, np.float32), np.ascontiguousarray(y, np.float32),
            np.ascontiguousarray(z, np.float32))


def homogeneous_media(shape, vp=6000.0, vs=3464.0, rho=2670.0):
    """md_gen_uniform_el_iso values (md_t.cu:396-413)."""
    mu = np.float32(vs * vs * rho)
    lam = np.float32(vp * vp * rho - 2.0 * float(mu))
    return (np.full(shape, lam, np.float32), np.full(shape, mu, np.float32), np.full(shape, np.float32(rho), np.float32))


def layered_media(z: np.ndarray, topo: np.ndarray, rng: np.random.Generator, perturb: float,
                  tops_km=(0.0, 1.0, 3.0, 6.0, 10.0), vs_kms=(1.8, 2.2, 2.7, 3.2, 3.6)):
    """5-layer 1-D model below the local surface: Vp = 1.73 Vs, Gardner
    density rho = 1741 Vp[km/s]^0.25, times (1 + perturb*N(0,1)) per node."""
    depth = (topo[None, :, :] - z) / 1000.0
    vs = np.full(z.shape, vs_kms[0], np.float64)
    for t, v in zip(tops_km, vs_kms):
        vs = np.where(depth >= t, v, vs)
    vp = 1.73 * vs
    rho = 1741.0 * vp ** 0.25
    vs *= 1000.0
    vp *= 1000.0
    if perturb > 0:
        f = 1.0 + perturb * rng.standard_normal(z.shape)
        vs = vs * f
        vp = vp * f
        rho = rho * f
    mu = vs * vs * rho
    lam = vp * vp * rho - 2.0 * mu
    return lam.astype(np.float32), mu.astype(np.float32), rho.astype(np.float32)


def von_karman(nj: int, nk: int, dh: float, corr: float, hurst: float, rng: np.random.Generator) -> np.ndarray:
    """Zero-mean, unit-std random field with a von Karman-like spectrum
    (FFT filtering of white Gaussian noise)."""
    noise = rng.standard_normal((nk, nj))
    ky = np.fft.fftfreq(nk, d=dh) * 2 * np.pi
    kx = np.fft.fftfreq(nj, d=dh) * 2 * np.pi
    kk = np.sqrt(ky[:, None] ** 2 + kx[None, :] ** 2)
    psd = (1.0 + (kk * corr) ** 2) ** (-(hurst + 1.0) / 2.0)
    f = np.real(np.fft.ifft2(np.fft.fft2(noise) * psd))
    f -= f.mean()
    s = f.std()
    return f / s if s > 0 else f


def _fault_fields(nj, nk):
    return {k: np.zeros((nk, nj), np.float32) for k in ("Tx", "Ty", "Tz", "mu_s", "mu_d", "Dc", "C0")}


def tpv_like(ni=80, nj=120, nk=60, dh=100.0, pml=12, fault_len_km=8.0, fault_dep_km=4.0,
             patch_km=1.5, hypo_depth_km=2.0, steps=800, dt=0.008, seed=1001, imethod=2,
             united_margin=None, nstations=16, free_top=True, strength_outside=1000.0,
             noise=0.02, dc=0.4, name="config1") -> Scenario:
    """Plain method, homogeneous Cartesian half-space with a vertical
    strike-slip fault from the surface (the 120x80x60 CPU-runnable case at
    its defaults)."""
    rng = _rng(seed)
    i0 = ni // 2
    x, y, z = cartesian_coords(ni, nj, nk, dh, i0)
    lam, mu, rho = homogeneous_media((nk, nj, ni))
    nlen = int(round(fault_len_km * 1000 / dh)) + 1
    ndep = int(round(fault_dep_km * 1000 / dh)) + 1
    j1 = (nj - nlen) // 2
    j2 = j1 + nlen - 1
    k2 = nk - 1 if free_top else nk - 1 - (nk - ndep) // 2
    k1 = k2 - ndep + 1
    f = _fault_fields(nj, nk)
    jj, kk = np.meshgrid(np.arange(nj), np.arange(nk))
    inside = (jj >= j1) & (jj <= j2) & (kk >= k1) & (kk <= k2)
    f["mu_s"][:] = np.where(inside, 0.68, strength_outside)
    f["mu_d"][:] = 0.52
    f["Dc"][:] = dc
    f["C0"][:] = 0.0
    f["Tx"][:] = -120.0 * MPA
    f["Ty"][:] = (70.0 * MPA * (1.0 + rng.uniform(-noise, noise, size=(nk, nj)))).astype(np.float32)
    npatch = int(round(patch_km * 1000 / dh)) + 1
    jc = (j1 + j2) // 2
    kc = (nk - 1) - int(round(hypo_depth_km * 1000 / dh)) if free_top else (k1 + k2) // 2
    pj1, pk1 = jc - npatch // 2, kc - npatch // 2
    patch = (jj >= pj1) & (jj < pj1 + npatch) & (kk >= pk1) & (kk < pk1 + npatch)
    f["Ty"][patch] = np.float32((0.68 * 120.0 + 0.5) * MPA)
    layers = [[pml, pml], [pml, pml], [pml, 0 if free_top else pml]]
    margin = united_margin if united_margin is not None else pml + 3
    st = None
    if nstations:
        si = rng.integers(pml + 1, ni - pml - 1, size=nstations)
        sj = rng.integers(pml + 1, nj - pml - 1, size=nstations)
        st = np.stack([si, sj, np.full(nstations, nk - 1)], axis=1).astype(np.int32)
    return Scenario(ni=ni, nj=nj, nk=nk, x=x, y=y, z=z, lam=lam, mu=mu, rho=rho, free_top=free_top,
                    pml_layers=layers, imethod=imethod,
                    faults=[FaultSpec(x_index=i0, grid=(j1, j2, k1, k2), fields=f)],
                    united_margin=margin, dt=dt, nt_total=steps + 1, stations=st, name=name)


def config1(**kw) -> Scenario:
    """120x80x60 at 100 m, 800 steps, dt 0.008 (reference axes 80x120x60)."""
    return tpv_like(**kw)


def small(ni=36, nj=48, nk=28, pml=5, imethod=2, free_top=True, curvilinear=False, seed=7,
          nfault=1, dc=0.1, **kw) -> Scenario:
    """A few-second oracle case with every boundary path active: PML on five
    faces, a free surface, one fault reaching the surface, optional
    topography."""
    dh = 100.0
    sc = tpv_like(ni=ni, nj=nj, nk=nk, dh=dh, pml=pml, fault_len_km=(nj - 2 * pml - 8) * dh / 1000.0,
                  fault_dep_km=(nk - pml - 6) * dh / 1000.0, patch_km=0.5, hypo_depth_km=0.9,
                  seed=seed, imethod=imethod, free_top=free_top, nstations=4, dc=dc, name="small", **kw)
    if curvilinear:
        rng = _rng(seed + 17)
        i0 = ni // 2
        xs = ((np.arange(ni) - i0) * dh)[None, :]
        ys = ((np.arange(nj) - nj // 2) * dh)[:, None]
        topo = hills(np.broadcast_to(xs, (nj, ni)), np.broadcast_to(ys, (nj, ni)), rng, 3, 150.0, 600.0, 1500.0)
        x, y, z = curvilinear_coords(ni, nj, nk, dh, i0, topo, dh, 1.3 * dh)
        sc.x, sc.y, sc.z = x, y, z
        lam, mu, rho = layered_media(z, topo, rng, 0.03, tops_km=(0.0, 0.5, 1.2, 1.8, 2.2))
        sc.lam, sc.mu, sc.rho = lam, mu, rho
        sc.dt = 0.006
        sc.name = "small_curv"
    if nfault == 2:
        # a second plane, parallel, further along x (welded everywhere except its own grid)
        f0 = sc.faults[0]
        f2 = {k: v.copy() for k, v in f0.fields.items()}
        sc.faults.append(type(f0)(x_index=f0.x_index + 8, grid=f0.grid, fields=f2))
    return sc


def config2(scale: float = 1.0, steps: int = 4000, seed: int = 2002, imethod: int = 2,
            united_margin: int | None = None) -> Scenario:
    """Plain method at one GPU's real shape: 1200x600x400 at 50 m (reference
    axes ni=600, nj=1200, nk=400), curvilinear with 6 Gaussian hills,
    vertically stretched (50 -> 75 m), layered medium with 5% noise, a
    40x18 km fault with von Karman shear stress, PML 20, dt 0.004.
    united_margin: the welded margin of the fault plane (fault_info.cu:
    1044-1060 hard-codes 40); by default PML layers + 3."""
    rng = _rng(seed)
    dh = 50.0
    ni, nj, nk = int(600 * scale), int(1200 * scale), int(400 * scale)
    pml = max(int(round(20 * scale)), 5)
    i0 = ni // 2
    xs = ((np.arange(ni) - i0) * dh)[None, :]
    ys = ((np.arange(nj) - nj // 2) * dh)[:, None]
    topo = hills(np.broadcast_to(xs, (nj, ni)), np.broadcast_to(ys, (nj, ni)), rng, 6, 800.0, 2000.0, 10000.0)
    x, y, z = curvilinear_coords(ni, nj, nk, dh, i0, topo, dh, 1.5 * dh)
    lam, mu, rho = layered_media(z, topo, rng, 0.05)
    nlen = min(int(round(40000 * scale / dh)) + 1, nj - 2 * (pml + 3) - 8)
    ndep = min(int(round(18000 * scale / dh)) + 1, nk - pml - 10)
    j1 = (nj - nlen) // 2
    j2 = j1 + nlen - 1
    k2 = nk - 1
    k1 = k2 - ndep + 1
    f = _fault_fields(nj, nk)
    # depth below the local surface along the fault plane (x = 0 column)
    depth_km = (topo[:, i0][None, :] - z[:, :, i0]) / 1000.0
    sn = -np.minimum(7.378 * depth_km, 120.0) * MPA
    sn = np.minimum(sn, -1.0 * MPA)          # keep the fault closed at the surface
    vk = von_karman(nj, nk, dh, 2000.0, 0.75, rng)
    tau = (0.6 + 0.05 * vk) * (-sn)
    mus = rng.uniform(0.62, 0.70, size=(nk, nj))
    jj, kk = np.meshgrid(np.arange(nj), np.arange(nk))
    inside = (jj >= j1) & (jj <= j2) & (kk >= k1) & (kk <= k2)
    f["mu_s"][:] = np.where(inside, mus, 1000.0)
    f["mu_d"][:] = mus - 0.15
    f["Dc"][:] = rng.uniform(0.3, 0.5, size=(nk, nj))
    f["Tx"][:] = sn
    f["Ty"][:] = tau
    npatch = int(round(2000 / dh)) + 1
    jc = (j1 + j2) // 2
    kc = (nk - 1) - int(round(min(9000.0 * scale, 0.5 * ndep * dh) / dh))
    patch = (jj >= jc - npatch // 2) & (jj < jc - npatch // 2 + npatch) & (kk >= kc - npatch // 2) & (kk < kc - npatch // 2 + npatch)
    f["Ty"][patch] = (mus[patch] * (-sn[patch]) + 1.0 * MPA)
    layers = [[pml, pml], [pml, pml], [pml, 0]]
    nst = 64
    si = rng.integers(pml + 1, ni - pml - 1, size=nst)
    sj = rng.integers(pml + 1, nj - pml - 1, size=nst)
    st = np.stack([si, sj, np.full(nst, nk - 1)], axis=1).astype(np.int32)
    return Scenario(ni=ni, nj=nj, nk=nk, x=x, y=y, z=z, lam=lam, mu=mu, rho=rho, free_top=True,
                    pml_layers=layers, imethod=imethod,
                    faults=[FaultSpec(x_index=i0, grid=(j1, j2, k1, k2), fields=f)],
                    united_margin=pml + 3 if united_margin is None else united_margin, dt=0.004,
                    nt_total=steps + 1, stations=st, name=f"config2x{scale:g}" if scale != 1.0 else "config2")


def config3(nscen: int = 32, scale: float = 1.0, steps_scale: float = 1.0, seed: int = 3000, imethod: int = 2,
            dims=(200, 400, 150), nhills: int = 3, hill_max: float = 800.0, name: str = "config3"):
    """Batched ensemble: ``nscen`` rupture scenarios on one 400x200x150 grid
    at 100 m (reference axes ni=200 fault-normal, nj=400 strike, nk=150),
    3 seeded Gaussian hills, layered medium as in config2, PML 20, dt 0.008.
    Scenario s (seed 3000+s) draws mu_s~U(0.62,0.70), mu_d=mu_s-0.15,
    Dc~U(0.3,0.5) m and a von Karman shear-stress field (mean 0.6 sigma_n,
    std 0.05 sigma_n), a fault length L~U(10,30) km set by locking the rest
    of the plane (mu_s = 1000), a 1-3 km nucleation patch at a random strike
    position and depth inside its fault, and a run length U(1000,3000) steps
    (times ``steps_scale``). Returns (base Scenario, [fault field dicts],
    [nt]); the base scenario carries scenario 0's fields and the widest
    fault grid."""
    rng = _rng(seed)
    dh = 100.0 / scale
    ni, nj, nk = int(dims[0] * scale), int(dims[1] * scale), int(dims[2] * scale)
    pml = max(int(round(20 * scale)), 5)
    i0 = ni // 2
    xs = ((np.arange(ni) - i0) * dh)[None, :]
    ys = ((np.arange(nj) - nj // 2) * dh)[:, None]
    topo = hills(np.broadcast_to(xs, (nj, ni)), np.broadcast_to(ys, (nj, ni)), rng, nhills, hill_max, 2000.0, 10000.0)
    x, y, z = curvilinear_coords(ni, nj, nk, dh, i0, topo, dh, 1.5 * dh)
    lam, mu, rho = layered_media(z, topo, rng, 0.05)
    nlen_max = min(int(round(30000 / dh)) + 1, nj - 2 * (pml + 3) - 8)
    ndep = min(int(round(18000 / dh)) + 1, nk - pml - 10)
    jc = nj // 2
    j1, j2 = jc - nlen_max // 2, jc - nlen_max // 2 + nlen_max - 1
    k2 = nk - 1
    k1 = k2 - ndep + 1
    depth_km = (topo[:, i0][None, :] - z[:, :, i0]) / 1000.0
    sn = np.minimum(-np.minimum(7.378 * depth_km, 120.0) * MPA, -1.0 * MPA)
    jj, kk = np.meshgrid(np.arange(nj), np.arange(nk))
    fields, nts = [], []
    for s in range(nscen):
        r = _rng(seed + s)
        f = _fault_fields(nj, nk)
        nlen = int(round(r.uniform(10000, 30000) / dh)) + 1
        nlen = min(nlen, nlen_max)
        a1 = jc - nlen // 2
        inside = (jj >= a1) & (jj < a1 + nlen) & (kk >= k1) & (kk <= k2)
        mus = r.uniform(0.62, 0.70, size=(nk, nj))
        f["mu_s"][:] = np.where(inside, mus, 1000.0)
        f["mu_d"][:] = mus - 0.15
        f["Dc"][:] = r.uniform(0.3, 0.5, size=(nk, nj))
        vk = von_karman(nj, nk, dh, 2000.0, 0.75, r)
        f["Tx"][:] = sn
        f["Ty"][:] = (0.6 + 0.05 * vk) * (-sn)
        npatch = int(round(r.uniform(1000, 3000) / dh)) + 1
        pj = int(r.integers(a1, a1 + max(nlen - npatch, 1)))
        pk = int(r.integers(k1, max(k2 - npatch, k1 + 1)))
        patch = (jj >= pj) & (jj < pj + npatch) & (kk >= pk) & (kk < pk + npatch) & inside
        f["Ty"][patch] = mus[patch] * (-sn[patch]) + 1.0 * MPA
        fields.append(f)
        nts.append(max(int(round(r.uniform(1000, 3000) * steps_scale)), 1))
    layers = [[pml, pml], [pml, pml], [pml, 0]]
    base = Scenario(ni=ni, nj=nj, nk=nk, x=x, y=y, z=z, lam=lam, mu=mu, rho=rho, free_top=True,
                    pml_layers=layers, imethod=imethod,
                    faults=[FaultSpec(x_index=i0, grid=(j1, j2, k1, k2), fields=fields[0])],
                    united_margin=pml + 3, dt=0.008 / scale, nt_total=nts[0], stations=None,
                    name=name if scale == 1.0 else f"{name}x{scale:g}")
    return base, fields, nts


def config5(nscen: int = 8, steps_scale: float = 1.0, seed: int = 5000, imethod: int = 2):
    """Batched adjoint ensemble: ``nscen`` scenarios of the config-3 physics
    (own friction and stress fields, fault lengths, patches) on a 320 x 200 x
    120 grid at 100 m (reference axes ni=200, nj=320, nk=120) with 4 seeded
    hills up to 500 m and PML 20."""
    return config3(nscen=nscen, steps_scale=steps_scale, seed=seed, imethod=imethod, dims=(200, 320, 120), nhills=4,
                   hill_max=500.0, name="config5")


def stress_drop(mode: str = "strike", imethod: int = 2, ni: int = 32, nj: int = 100, nk: int = 60, dh: float = 100.0,
                pml: int = 5, sigma_n: float = 100.0 * MPA, mu_d: float = 0.5, dtau: float = 10.0 * MPA,
                dt: float = 0.008, steps: int = 80) -> Scenario:
    """Uniform stress drop on a fault covering the plane: a homogeneous
    Cartesian medium, mu_s = mu_d, and an initial shear traction dtau above
    the dynamic strength everywhere, so every unwelded node slips from the
    first step. Until waves from the welded border arrive, the solution is
    the 1-D one: slip rate 2 dtau / (rho Vs), side velocities +-dtau / (rho
    Vs), shear traction mu_d sigma_n, slip = rate * t.

    mode "strike": shear along y (strike) with a free surface on top (that
    solution has sigma_yz = sigma_zz = 0, so it is traction-free at the
    surface); mode "dip": shear along z with the fault buried (PML on all six
    faces)."""
    if mode not in ("strike", "dip"):
        raise ValueError("mode must be 'strike' or 'dip'")
    free_top = mode == "strike"
    i0 = ni // 2
    x, y, z = cartesian_coords(ni, nj, nk, dh, i0)
    lam, mu, rho = homogeneous_media((nk, nj, ni))
    f = _fault_fields(nj, nk)
    f["mu_s"][:] = mu_d
    f["mu_d"][:] = mu_d
    f["Dc"][:] = 0.4
    f["C0"][:] = 0.0
    f["Tx"][:] = -sigma_n
    f[("Ty" if mode == "strike" else "Tz")][:] = mu_d * sigma_n + dtau
    layers = [[pml, pml], [pml, pml], [pml, 0 if free_top else pml]]
    return Scenario(ni=ni, nj=nj, nk=nk, x=x, y=y, z=z, lam=lam, mu=mu, rho=rho, free_top=free_top,
                    pml_layers=layers, imethod=imethod,
                    faults=[FaultSpec(x_index=i0, grid=(0, nj - 1, 0, nk - 1), fields=f)],
                    united_margin=pml + 3, dt=dt, nt_total=steps + 1, stations=None,
                    name=f"stress_drop_{mode}")
