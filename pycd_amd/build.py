"""Build libpycd_hip.so for gfx950 in-tree (pycd_amd/lib/).

    python -m pycd_amd.build

Compiles every pycd_amd/csrc/*.hip with hipcc (one object per file, in
parallel) and links the shared library next to this package, where
`pycd_amd._native` loads it from.
"""
import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "csrc"
LIB_DIR = HERE / "lib"
LIB = LIB_DIR / "libpycd_hip.so"
OBJ_DIR = HERE / "_build"
ARCH = os.environ.get("PYCD_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-munsafe-fp-atomics",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(HERE.parent / "include"),
]


def _hipcc():
    root = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(root) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def _compile(src, extra=(), obj_dir=None):
    obj = (obj_dir or OBJ_DIR) / (src.stem + ".o")
    cmd = [_hipcc(), *HIP_FLAGS, *extra, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr}")
    return obj


def build(verbose=True, out=None, defines=(), obj_dir=None):
    """Compile and link; `defines` are extra -D flags (tuning variants are
    built to `out` with their own `obj_dir`)."""
    obj_dir = Path(obj_dir) if obj_dir else OBJ_DIR
    out = Path(out) if out else LIB
    obj_dir.mkdir(parents=True, exist_ok=True)
    out.parent.mkdir(parents=True, exist_ok=True)
    sources = sorted(SRC.glob("*.hip"))
    extra = [f"-D{d}" for d in defines]
    workers = min(len(sources), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        objs = list(pool.map(lambda s: _compile(s, extra, obj_dir), sources))
    tmp = out.with_suffix(".so.tmp")
    cmd = [_hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC",
           *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out} from {len(sources)} sources")
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
