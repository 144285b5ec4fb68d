"""Build recipe for libfluxhip.so (gfx950).

Compiles every ``csrc/*.hip`` with hipcc into one shared library next to this
file.  No torch extension machinery: the library is a plain C-ABI loaded with
ctypes, so it travels with the source tree.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libfluxhip.so"
ARCH = os.environ.get("FLUXHIP_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed fp32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    # a value-returning function that can fall off its end is an error, not a warning
    "-Werror=return-type",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; libfluxhip needs ROCm to build")
    return path


def build(verbose=False, jobs=None):
    """Compile all sources (always recompiles) and link libfluxhip.so."""
    sources = sorted(CSRC.glob("*.hip"))
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()

    def compile_one(src):
        obj = objdir / (src.stem + ".o")
        cmd = [cc, *FLAGS, "-c", str(src), "-o", str(obj)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr[-6000:]}")
        if verbose and res.stderr.strip():
            print(res.stderr, file=sys.stderr)
        return obj

    jobs = jobs or min(8, len(sources)) or 1
    with ThreadPoolExecutor(jobs) as pool:
        objs = list(pool.map(compile_one, sources))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-4000:]}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
