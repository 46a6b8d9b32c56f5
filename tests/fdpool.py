"""Runs independent FP64 oracle evaluations (the terms of central
differences) in a pool of CPU worker processes, each with a share of the
OpenMP threads. The workers import only the oracle and the scenario
generators: nothing in them touches the GPU."""
from __future__ import annotations

import os
from concurrent.futures import ProcessPoolExecutor, as_completed
import multiprocessing as mp
import time


def _init(threads: int, paths):
    import sys
    for p in paths:
        if p not in sys.path:
            sys.path.insert(0, p)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import oracle
    oracle.Oracle.threads(threads)


def cpu_share() -> int:
    """Cores this process may use: its affinity set, OMP_NUM_THREADS and at
    most 16 (the host share of one GPU's job on an MI355X node)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, min(n, 16))


def run_all(fn, jobs, workers: int | None = None, label: str = "fp64"):
    """[fn(*job) for job in jobs], evaluated in parallel worker processes,
    with a progress line at least every 30 s."""
    import sys
    from helpers import progress
    cores = cpu_share()
    workers = max(1, min(workers or cores, len(jobs), cores))
    threads = max(1, cores // workers)
    ctx = mp.get_context("spawn")
    out = [None] * len(jobs)
    t0 = last = time.time()
    with ProcessPoolExecutor(max_workers=workers, mp_context=ctx, initializer=_init,
                             initargs=(threads, list(sys.path))) as ex:
        futs = {ex.submit(fn, *job): n for n, job in enumerate(jobs)}
        done = 0
        for f in as_completed(futs):
            out[futs[f]] = f.result()
            done += 1
            if time.time() - last > 30 or done == len(jobs):
                last = time.time()
                progress(f"[{label}] {done}/{len(jobs)} evaluations, {last - t0:.0f} s")
    return out
