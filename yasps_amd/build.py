"""Build recipe of the yasps_amd HIP library (gfx950).

``python -m yasps_amd.build`` compiles every ``csrc/*.hip`` with hipcc and links
``yasps_amd/libyasps_amd.so`` in-tree (the shared object travels with the source tree).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libyasps_amd.so"
ARCH = os.environ.get("YSP_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    "-I", str(PKG.parent / "include"),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the yasps_amd HIP library cannot be built")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def hip_runtime_dir() -> str:
    """Directory of the HIP runtime the process uses. PyTorch-ROCm ships its own
    libamdhip64.so; the library links that one so that torch and the kernels share one
    runtime (streams, allocations and ordering are per runtime)."""
    try:
        import torch  # noqa: F401
        d = Path(torch.__file__).resolve().parent / "lib"
        if (d / "libamdhip64.so").exists():
            return str(d)
    except Exception:
        pass
    return "/opt/rocm/lib"


def build(verbose: bool = True, only=None) -> Path:
    """Compile every source (or only the named stems, reusing the other objects from an
    earlier build) and link the library."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    todo = sources if not only else [s for s in sources if s.stem in set(only)]
    if only and len(todo) != len(set(only)):
        raise ValueError(f"unknown sources in {only}")
    workers = max(1, min(len(todo), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(_compile, todo))
    objs = [BUILD / (s.stem + ".o") for s in sources]
    missing = [o.name for o in objs if not o.exists()]
    if missing:
        raise RuntimeError(f"objects missing (build everything first): {missing}")
    tmp = LIB.with_suffix(".so.tmp")
    rt = hip_runtime_dir()
    # plain host link: device code is already bundled in every object
    cmd = ["g++", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), f"-L{rt}",
           "-l:libamdhip64.so", f"-Wl,-rpath,{rt}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources")
    return LIB


if __name__ == "__main__":
    build(only=sys.argv[1:] or None)
    sys.exit(0)
