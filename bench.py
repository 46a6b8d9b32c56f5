#!/usr/bin/env python3
"""Throughput of the time-stepping core on one MI355X.

Workload: the plain-method configuration at one GPU's real shape,
1200 x 600 x 400 nodes (reference axes ni=600 fault-normal, nj=1200 strike,
nk=400 depth) at 50 m, curvilinear with topography, layered medium, one
40 x 18 km slip-weakening fault, CFS-PML on five faces and a free surface.
A step is one full RK4 time step (4 stages) of the whole core: fault plane,
volume stencil with PML and free surface, split-node columns, accumulators.
Inputs are resident in HBM before the timed region.

Prints one JSON line (see the keys below). ``--gpus N`` with N > 1 exits:
this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "gpu-cgfd3d-dynamic_amd"))
sys.path.insert(0, str(REPO))

METRIC = "Gcell-updates/s + %HBM peak; batched vs plain; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
# The reference's RK4 bookkeeping touches, per point and stage (FP32): stage 0
# reads cur(=pre) 36 + metric 36 + media 12 and writes tmp 36 + end 36;
# stages 1-2 also read pre 36 + end 36; stage 3 reads cur 36 + 48 + end 36 and
# writes end 36: 768 B per cell-update, the figure the HBM fraction of the
# whole step is quoted on (hbm_frac_768B_model).
BYTES_PER_CELL_UPDATE = 768
# The arrays one launch of this volume update reads and writes per point
# (reads, writes): the stage input, the metric (five terms on a grid of
# vertical columns, else nine), three media arrays and the levels the stage
# combines (stages 1-2: the step's start; stage 3: the start and the outputs
# of stages 0 and 1); one output level. End accumulation is deferred to
# stage 3, so stages 0-2 write one level.
STAGE_ARRAYS = {True: ((17, 9), (26, 9), (26, 9), (44, 9)), False: ((21, 9), (30, 9), (30, 9), (48, 9))}


def stage_bytes(vcol: bool):
    return tuple(4 * (r + w) for r, w in STAGE_ARRAYS[vcol])  # 596 or 660 per cell-update


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def read_traffic(profile_dir: Path):
    """HBM bytes per volume launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE in KiB, scaled by the copy-kernel calibration)."""
    f = profile_dir / "k_vol_config2_pmc.txt"
    if not f.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def calib_copy(mbytes: int, reps: int = 1) -> float:
    """Device-to-device copies of a known size: the reference point that
    scales the FETCH_SIZE/WRITE_SIZE counters of a profiling run, and, timed
    with HIP events, the streaming rate a plain copy reaches on this GPU
    (read + written bytes per second, GB/s)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = C.c_size_t(mbytes << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), n) or hip.hipMalloc(C.byref(dst), n):
        raise RuntimeError("hipMalloc failed in calib_copy")
    hip.hipMemset(src, 0, n)
    hip.hipDeviceSynchronize()
    e0, e1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    hip.hipMemcpy(dst, src, n, 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipEventRecord(e0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(dst, src, n, 3, None)
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), e0, e1)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(src)
    hip.hipFree(dst)
    gbs = 2.0 * (mbytes << 20) * reps / (ms.value / 1e3) / 1e9 if ms.value > 0 else 0.0
    log(f"[bench] device copy of {mbytes} MiB x {reps}: {gbs:.0f} GB/s (read + write)")
    return gbs


def stream_ceiling(ni: int, nj: int, nk: int, vcol: bool, reps: int = 5):
    """(GB/s, per-stage GB/s): a kernel that streams each RK stage's arrays
    per point (STAGE_ARRAYS) over the same grid with no stencil
    (cgfd_calib_stream); the step's rate is its bytes over the sum of the
    stages' times at those rates."""
    import ctypes as C
    from cgfd import lib as L
    lib = L.load_library()
    rates = {}
    for r, w in set(STAGE_ARRAYS[vcol]):
        g = C.c_float()
        L.check(lib, lib.cgfd_calib_stream(0, ni, nj, nk, r, w, reps, C.byref(g)))
        rates[(r, w)] = float(g.value)
    per_stage = [rates[rw] for rw in STAGE_ARRAYS[vcol]]
    byts = stage_bytes(vcol)
    ceiling = sum(byts) / sum(b / g for b, g in zip(byts, per_stage))
    log(f"[bench] the stages' data movement without the stencil, {ni}x{nj}x{nk}: "
        f"{', '.join(f'{g:.0f}' for g in per_stage)} GB/s; step {ceiling:.0f} GB/s")
    return ceiling, per_stage


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(threads: int, steps: int = 24):
    """The oracle (FP32 CPU restatement, OpenMP over k-planes) on a bounded
    sample: the same configuration generator at 1/4 linear scale (150 x 300
    x 100 nodes). Timed with the -O3 AVX2/FMA build of the restatement when
    this CPU has those instructions (the checker build is -O2 without
    contraction)."""
    from cgfd import configs, model
    import oracle
    fast = oracle.load_fast()
    sc = configs.config2(scale=0.25, steps=steps)
    st = model.Setup(sc)
    n = oracle.Oracle.threads(threads)
    o = oracle.Oracle(st.desc, sc.shape_ext, len(sc.faults))
    o.run(1)
    t = time.perf_counter()
    o.run(steps)
    el = time.perf_counter() - t
    cells = sc.ni * sc.nj * sc.nk * steps
    o.close()
    st.close()
    build = oracle.active_build()
    return {"value": cells / el / 1e9, "unit": "Gcell-updates/s", "cores": n, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": _cpu_model(), "build": build,
            "sample": f"oracle FP32 ({build}) on the same generator at 1/4 scale ({sc.ni}x{sc.nj}x{sc.nk}), "
                      f"{steps} steps after 1 warm-up, {n} OpenMP threads of {os.cpu_count()} visible, {el:.1f} s"
                      + ("" if fast else "; the CPU lacks AVX2/FMA, so the portable build ran")}


def batched_leg(steps: int, nscen: int, variant: int, overlap: bool = False):
    """Config-3 ensemble: nscen scenarios on one 400x200x150 grid, batched,
    against scenario 0 alone through the plain path on the same grid, both
    with the same volume-kernel variant (0: autotuned for each). Two timed
    batched runs: every scenario active for all the timed steps, and the
    config's uneven run lengths scaled into the timed window (U(1000,3000)
    steps mapped to U(steps/3, steps)), so scenarios retire inside it; the
    second counts only the scenario-steps actually taken."""
    from cgfd import configs, model
    t0 = time.perf_counter()
    base, fields, nts = configs.config3(nscen=nscen)
    st = model.Setup(base)
    log(f"[bench] config3 {base.ni}x{base.nj}x{base.nk}, {nscen} scenarios, setup {time.perf_counter() - t0:.1f} s")
    cells = base.ni * base.nj * base.nk
    # plain: scenario 0 alone
    core = model.Core(st, variant=variant)
    if variant == 0:
        core.autotune(2)
    vplain = core.variant_info()
    core.set_overlap(overlap)
    core.run(8)
    core.sync()
    t = time.perf_counter()
    core.run(steps)
    core.sync()
    plain = cells * steps / (time.perf_counter() - t) / 1e9
    core.close()
    # batched, every scenario active for the timed steps
    big = [n + steps + 16 for n in nts]
    bat = model.Batch(st, [(None if s == 0 else [fields[s]], big[s]) for s in range(nscen)], variant=variant)
    if variant == 0:
        bat.autotune(2)
    vbat = bat.variant_info()
    bat.set_overlap(overlap)
    bat.run(8)
    bat.sync()
    t = time.perf_counter()
    bat.run(steps)
    bat.sync()
    el = time.perf_counter() - t
    batched = cells * nscen * steps / el / 1e9
    bat.close()
    # batched, uneven run lengths retiring inside the timed window
    lo, hi = 1000, 3000
    short = [8 + max(1, int(round(steps * (1 + 2 * (n - lo) / (hi - lo)) / 3))) for n in nts]
    bat = model.Batch(st, [(None if s == 0 else [fields[s]], short[s]) for s in range(nscen)], variant=variant)
    if variant != 0:
        bat.set_variant(variant)
    else:
        bat.set_variant(vbat[0])
        bat.set_zchunk(vbat[1])
    bat.run(8)
    bat.sync()
    t = time.perf_counter()
    bat.run(max(short) - 8)
    bat.sync()
    el_u = time.perf_counter() - t
    taken = sum(n - 8 for n in short)
    uneven = cells * taken / el_u / 1e9
    bat.close()
    st.close()
    return {"value": round(batched, 4), "unit": "scenario-Gcell-updates/s", "nscen": nscen,
            "grid": [base.ni, base.nj, base.nk], "steps": steps, "ms_per_step": round(el / steps * 1e3, 3),
            "plain_same_grid": round(plain, 4), "per_scenario_vs_plain": round(batched / plain, 3),
            "volume_kernel": {"plain": list(vplain), "batched": list(vbat), "autotuned": variant == 0},
            "uneven": {"value": round(uneven, 4), "per_scenario_vs_plain": round(uneven / plain, 3),
                       "timed_steps": max(short) - 8, "scenario_steps": taken,
                       "run_lengths": f"{min(short) - 8}..{max(short) - 8} timed steps per scenario"},
            "note": "value: all scenarios active in the timed steps; uneven: scenarios retire inside the timed "
                    "window, counting only the scenario-steps taken"}


def adjoint_leg(steps: int):
    """Adjoint/forward time on the differentiation grid (160 x 240 x 100
    nodes at 100 m, homogeneous half-space with 3% noise, free surface,
    PML 12, a 12 x 6 km fault from the surface rupturing from its nucleation
    patch, 32 surface stations). Forward: `steps` steps recording the
    stations and the friction tape; adjoint: the gradient of 1/2 |traces|^2
    with respect to the initial wavefield and to mu_s, mu_d, Dc, C0 and T0
    at every fault node, over the same steps."""
    import numpy as np
    from cgfd import configs, model
    sc = configs.tpv_like(ni=160, nj=240, nk=100, dh=100.0, pml=12, fault_len_km=12.0, fault_dep_km=6.0,
                          nstations=32, seed=4004, name="config4")
    rng = np.random.default_rng(4004)
    sc.mu = (sc.mu * (1.0 + 0.03 * rng.standard_normal(sc.mu.shape))).astype(np.float32)
    st = model.Setup(sc)
    core = model.Core(st)
    core.set_stations(sc.stations + 3, steps)
    core.adjoint_forward(steps)
    core.sync()
    t = time.perf_counter()
    core.run(steps)
    core.sync()
    t_fwd = time.perf_counter() - t
    tr = core.read_stations()
    t = time.perf_counter()
    g, gf = core.adjoint_fault(tr)
    t_adj = time.perf_counter() - t
    core.close()
    st.close()
    return {"value": round(t_adj / t_fwd, 3),
            "scope": "whole time loop: interior, CFS-PML, free surface, split-node fault with slip-weakening friction",
            "grid": [sc.ni, sc.nj, sc.nk], "steps": steps, "forward_ms_per_step": round(t_fwd / steps * 1e3, 3),
            "adjoint_ms_per_step": round(t_adj / steps * 1e3, 3),
            "finite": bool(np.isfinite(g).all() and np.isfinite(gf).all()),
            "gradients": "initial wavefield; mu_s, mu_d, Dc, C0, T0x, T0y, T0z per fault node"}


def batched_adjoint_leg(steps: int, nscen: int = 8):
    """Batched forward (per-scenario friction tapes) and per-scenario
    adjoints of nscen rupture scenarios on the 320 x 200 x 120 grid (4 hills,
    PML 20, own friction/stress fields), 32 surface stations, against
    scenario 0 run and differentiated alone on the same grid."""
    import numpy as np
    from cgfd import configs, model
    base, fields, _ = configs.config5(nscen=nscen)
    rng = np.random.default_rng(5000)
    base.stations = np.stack([rng.integers(23, base.ni - 23, 32), rng.integers(23, base.nj - 23, 32),
                              np.full(32, base.nk - 1)], 1).astype(np.int32)
    st = model.Setup(base)
    bat = model.Batch(st, [(None if s == 0 else [fields[s]], steps + 5) for s in range(nscen)])
    bat.set_stations(base.stations + 3, steps)
    bat.adjoint_forward(steps)
    bat.sync()
    t = time.perf_counter()
    bat.run(steps)
    bat.sync()
    t_bf = time.perf_counter() - t
    t = time.perf_counter()
    finite = True
    for s in range(nscen):
        g, gf = bat.adjoint_fault(s, bat.read_stations(s))
        finite = finite and bool(np.isfinite(gf).all())
    t_ba = time.perf_counter() - t
    bat.close()
    core = model.Core(st)
    core.set_stations(base.stations + 3, steps)
    core.adjoint_forward(steps)
    core.sync()
    t = time.perf_counter()
    core.run(steps)
    core.sync()
    t_f = time.perf_counter() - t
    t = time.perf_counter()
    core.adjoint_fault(core.read_stations())
    t_a = time.perf_counter() - t
    core.close()
    st.close()
    per = (t_bf + t_ba) / nscen
    return {"nscen": nscen, "grid": [base.ni, base.nj, base.nk], "steps": steps,
            "batched_forward_s_per_scenario": round(t_bf / nscen, 3), "adjoint_s_per_scenario": round(t_ba / nscen, 3),
            "adjoint_vs_batched_forward": round(t_ba / t_bf, 3),
            "unbatched_forward_s": round(t_f, 3), "unbatched_adjoint_s": round(t_a, 3),
            "per_scenario_forward_plus_adjoint_vs_unbatched": round(per / (t_f + t_a), 3), "finite": finite}


def device_setup_leg(sc, st_host, host_s):
    """The setup again with the metric and the fault coefficients computed
    on the device (csrc/k_setup.hip), timed end to end (host arrays in, host
    arrays out, as the descriptor needs them), and its metric compared with
    the host setup's bit for bit."""
    import ctypes as C
    import numpy as np
    from cgfd import model
    t0 = time.perf_counter()
    st_d = model.Setup(sc, device_setup=True)
    dev_s = time.perf_counter() - t0
    nz, ny, nx = sc.shape_ext
    n = 10 * nz * ny * nx

    def view(p):
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n,))
    same = bool(np.array_equal(view(st_host.desc.metric), view(st_d.desc.metric)))
    st_d.close()
    log(f"[bench] device setup in {dev_s:.1f} s (host {host_s:.1f} s), metric identical: {same}")
    return {"host_s": round(host_s, 2), "device_s": round(dev_s, 2), "metric_bit_identical": same,
            "host_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
            "scope": "metric (gd_curv_metric_cal) and fault coefficients (fault_coef_cal) on the device; "
                     "coordinates, media, PML profiles and fault_set on the host; times include the "
                     "host<->device copies"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--scale", type=float, default=1.0, help="linear scale of the workload (1.0 = 1200x600x400)")
    ap.add_argument("--variant", type=int, default=0,
                    help="volume kernel variant (0 = autotune and use the fastest; see cgfd_core_set_variant)")
    ap.add_argument("--zchunk", type=int, default=0, help="y tiles per slab (0 = default or autotuned)")
    ap.add_argument("--overlap", action="store_true",
                    help="stage fork: interior, PML + top and fault kernels on three streams")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched leg (profiling)")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint leg (profiling)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--batched-steps", type=int, default=16)
    ap.add_argument("--nscen", type=int, default=32)
    ap.add_argument("--batched-variant", type=int, default=0)
    ap.add_argument("--no-copy-ceiling", action="store_true")
    ap.add_argument("--no-setup-leg", action="store_true")
    ap.add_argument("--calib-copy", type=int, default=0, help="MiB of a device copy to run first (profiling)")
    ap.add_argument("--united-margin", type=int, default=None,
                    help="welded margin of the fault plane (default: PML layers + 3; the reference hard-codes 40)")
    ap.add_argument("--cpu-threads", type=int, default=min(os.cpu_count() or 1, 16))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs the whole domain on one GPU; --gpus {args.gpus} is not supported "
              "(sharding is design-only)", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    if args.calib_copy:
        calib_copy(args.calib_copy)
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_leg(args.adjoint_steps)}), flush=True)
        return 0
    if args.batched_only:
        print(json.dumps({"batched": batched_leg(args.batched_steps, args.nscen, args.batched_variant,
                                                 args.overlap)}), flush=True)
        return 0
    from cgfd import configs, model
    t0 = time.perf_counter()
    sc = configs.config2(scale=args.scale, steps=args.steps + args.warmup, united_margin=args.united_margin)
    log(f"[bench] generated {sc.name} {sc.ni}x{sc.nj}x{sc.nk} in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    st = model.Setup(sc)
    host_setup_s = time.perf_counter() - t0
    log(f"[bench] host setup in {host_setup_s:.1f} s")
    setup_leg = None
    if not args.no_setup_leg:
        setup_leg = device_setup_leg(sc, st, host_setup_s)
    t0 = time.perf_counter()
    core = model.Core(st, variant=args.variant)
    if args.zchunk:
        core.set_zchunk(args.zchunk)
    # the 64 stations recorded at every step, as the reference's io_recv_keep
    core.set_stations(sc.stations + 3, args.steps + args.warmup)
    if args.overlap:
        core.set_overlap(True)
    core.sync()
    log(f"[bench] device upload in {time.perf_counter() - t0:.1f} s")
    tuned = None
    if args.variant == 0:
        t0 = time.perf_counter()
        tuned = core.autotune(2)
        log(f"[bench] autotune: variant {tuned[0]}, zchunk {tuned[1]}, {tuned[2]:.3f} ms per stage-0 launch "
            f"({time.perf_counter() - t0:.1f} s)")
    # free host copies (the core holds everything in HBM)
    del sc.x, sc.y, sc.z, sc.lam, sc.mu, sc.rho
    if args.warmup:
        core.run(args.warmup)
        core.sync()
    log(f"[bench] warm-up {args.warmup} steps done")
    ms_total, ms_vol, n_vol = core.run_timed(args.steps)
    core.sync()
    log(f"[bench] timed {args.steps} steps: {ms_total:.1f} ms total, volume {ms_vol:.1f} ms in {n_vol} launches")
    w = core.read_fault(0)
    import numpy as np
    slip_max = float(np.nanmax(w[6]))
    vcol_grid, vcol_used = core.vertical_columns()
    sbytes = stage_bytes(vcol_used)

    ni, nj, nk = sc.ni, sc.nj, sc.nk
    cells = ni * nj * nk * args.steps
    sec = ms_total / 1e3
    value = cells / sec / 1e9
    hbm_frac = value * 1e9 * BYTES_PER_CELL_UPDATE / (HBM_PEAK_GBS * 1e9)
    nvol_planes = nk - 4  # the top four layers go to the free-surface kernel
    vol_pts = ni * nj * nvol_planes
    launches_per_step = 4
    bytes_vol = sum(sbytes) * vol_pts * (n_vol // launches_per_step)
    achieved = bytes_vol / (ms_vol / 1e3) / 1e9 if ms_vol > 0 else 0.0
    traffic = read_traffic(REPO / "profiles")
    copy_gbs = None if args.no_copy_ceiling else calib_copy(2048, reps=5)
    core.close()
    st.close()
    stream_gbs, stream_stages = (None, None) if args.no_copy_ceiling else stream_ceiling(ni, nj, nk, vcol_used)
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"{sc.name}: plain-method rupture, {nj}x{ni}x{nk} nodes (strike x normal x depth) "
                               f"at 50 m, curvilinear, PML 20, free surface, 1 fault, FP32",
                   "grid": [ni, nj, nk], "parallelism": "single GPU", "hbm_frac_768B_model": round(hbm_frac, 4),
                   "max_slip_m": round(slip_max, 4), "stage_fork": args.overlap, "stations": int(len(sc.stations)),
                   "united_margin": int(sc.united_margin),
                   "volume_kernel": {"variant": tuned[0] if tuned else args.variant,
                                     "zchunk": tuned[1] if tuned else args.zchunk, "autotuned": tuned is not None}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_vol (volume stencil + PML + RK update)",
                     "bytes_per_cell_update": sum(sbytes),
                     "metric_form": ("five terms (vertical columns: xi_y, xi_z, eta_x, eta_z below 2^-32 of "
                                     "xi_x, eta_y)" if vcol_used else "nine terms"),
                     "avg_launch_ms": round(ms_vol / max(n_vol, 1), 4),
                     "traffic_source": "profiles/k_vol_config2_pmc.txt: rocprofv3 FETCH_SIZE + WRITE_SIZE of the "
                                       "volume dispatches, calibrated with a device copy (not measured in this run)",
                     "copy_gbs": round(copy_gbs, 1) if copy_gbs else None,
                     "frac_of_copy": round(achieved / copy_gbs, 4) if copy_gbs else None,
                     "stream_ceiling_gbs": round(stream_gbs, 1) if stream_gbs else None,
                     "frac_of_stream_ceiling": round(achieved / stream_gbs, 4) if stream_gbs else None,
                     "stream_ceiling_stages_gbs": [round(g, 1) for g in stream_stages] if stream_stages else None,
                     "stream_ceiling_note": "cgfd_calib_stream: each RK stage's reads + writes per point "
                                            f"{list(STAGE_ARRAYS[vcol_used])} with no stencil, same grid and "
                                            "tiles, measured in this run; the step's rate is its bytes over the "
                                            "stages' times at those rates"},
    }
    if setup_leg:
        out["setup"] = setup_leg
    del core, st, sc
    if not args.no_batched:
        log("[bench] batched ensemble ...")
        out["batched"] = batched_leg(args.batched_steps, args.nscen, args.batched_variant, args.overlap)
    out["adjoint_vs_forward"] = None
    if not args.no_adjoint:
        log("[bench] adjoint ...")
        out["adjoint"] = adjoint_leg(args.adjoint_steps)
        out["adjoint_vs_forward"] = out["adjoint"]["value"]
        log("[bench] batched adjoint ...")
        out["batched_adjoint"] = batched_adjoint_leg(args.adjoint_steps)
    if rank == 0 and not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_threads)
    if rank == 0:
        print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
