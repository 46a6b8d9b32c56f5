"""Build recipes for the native libraries (in-tree, no build system).

``build_core()``   hipcc -> gpu-cgfd3d-dynamic_amd/_build/libcgfd.so
                   (HIP kernels for gfx950 + the C ABI + the host setup)
``build_oracle()`` g++   -> oracle/_build/libcgfd_oracle.so (test infrastructure)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent.parent          # gpu-cgfd3d-dynamic_amd/
REPO = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
HOST = PKG_DIR / "host"
INCLUDE = REPO / "include"
BUILD_DIR = PKG_DIR / "_build"
LIB_PATH = BUILD_DIR / "libcgfd.so"
ORACLE_DIR = REPO / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "libcgfd_oracle.so"
ORACLE_FAST_LIB = ORACLE_DIR / "_build" / "libcgfd_oracle_fast.so"

HIP_SOURCES = ["core.hip", "k_volume.hip", "k_fault.hip", "k_adjoint.hip", "k_adj_fault.hip", "k_io.hip", "k_setup.hip",
               "k_calib.hip"]
HOST_SOURCES = ["cgfd_setup.cpp", "cgfd_errors.cpp", "cgfd_nc.cpp", "cgfd_sac.cpp", "cgfd_stations.cpp"]

# Packed FP32 instructions stay disabled: hipcc otherwise forms v_pk_*_f32
# from scalar float code.
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result", "-fopenmp",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libcgfd.so)")


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def _sources() -> list[Path]:
    return [CSRC / s for s in HIP_SOURCES if (CSRC / s).exists()] + [HOST / s for s in HOST_SOURCES]


def _digest(paths) -> str:
    import hashlib
    h = hashlib.sha256()
    for p in sorted(paths):
        h.update(p.name.encode())
        h.update(p.read_bytes())
    return h.hexdigest()


def _core_deps() -> list[Path]:
    return _sources() + list(CSRC.glob("*.h")) + list(HOST.glob("*.h")) + list(INCLUDE.glob("*.h"))


def _oracle_deps() -> list[Path]:
    return list(ORACLE_DIR.glob("*.cpp")) + list(INCLUDE.glob("*.h"))


def _stale(lib: Path, deps) -> bool:
    """A library is stale when it is missing or when the digest of the
    sources it was built from (stamp file beside it) differs from the tree's:
    content, not file times, so a copied tree with new mtimes is not stale."""
    stamp = lib.with_suffix(".stamp")
    if not lib.exists() or not stamp.exists():
        return True
    return stamp.read_text().strip() != _digest(deps)


def core_is_stale() -> bool:
    return _stale(LIB_PATH, _core_deps())


def oracle_is_stale() -> bool:
    return _stale(ORACLE_LIB, _oracle_deps())


def build_core(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile the HIP core + host setup into one shared library: one object
    per source, compiled in parallel, then one link."""
    from concurrent.futures import ThreadPoolExecutor
    BUILD_DIR.mkdir(parents=True, exist_ok=True)
    srcs = _sources()
    flags = [f for f in HIP_FLAGS if f != "-shared"]
    inc = [f"-I{INCLUDE}", f"-I{CSRC}", f"-I{HOST}"]
    objs = [BUILD_DIR / (s.name + ".o") for s in srcs]
    cmds = [[_hipcc(), *flags, *inc, "-c", str(s), "-o", str(o)] for s, o in zip(srcs, objs)]
    if verbose:
        for c in cmds:
            print(" ".join(c))
    n = jobs or min(len(cmds), max(1, min(os.cpu_count() or 1, 8)))
    with ThreadPoolExecutor(max_workers=n) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB_PATH.with_suffix(".so.tmp")
    link = [_hipcc(), *HIP_FLAGS, *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(link))
    digest = _digest(_core_deps())
    _run(link)
    os.replace(tmp, LIB_PATH)
    LIB_PATH.with_suffix(".stamp").write_text(digest + "\n")
    return LIB_PATH


def build_oracle(verbose: bool = False) -> Path:
    """Compile the CPU restatement: the checker (OpenMP, -O2, no FMA
    contraction) and the same source -O3 for AVX2/FMA, which bench.py times
    as the CPU baseline."""
    from concurrent.futures import ThreadPoolExecutor
    out = ORACLE_LIB
    out.parent.mkdir(parents=True, exist_ok=True)
    src = str(ORACLE_DIR / "cgfd_oracle.cpp")
    tmp = out.with_suffix(".so.tmp")
    tmp_fast = ORACLE_FAST_LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-march=x86-64-v2",
           "-ffp-contract=off", f"-I{INCLUDE}", src, "-o", str(tmp)]
    cmd_fast = ["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-march=x86-64-v3",
                f"-I{INCLUDE}", src, "-o", str(tmp_fast)]
    if verbose:
        print(" ".join(cmd))
        print(" ".join(cmd_fast))
    digest = _digest(_oracle_deps())
    with ThreadPoolExecutor(max_workers=2) as ex:
        for f in [ex.submit(_run, c) for c in (cmd, cmd_fast)]:
            f.result()
    os.replace(tmp, out)
    os.replace(tmp_fast, ORACLE_FAST_LIB)
    out.with_suffix(".stamp").write_text(digest + "\n")
    return out


if __name__ == "__main__":
    print(build_core(verbose=True))
    print(build_oracle(verbose=True))
