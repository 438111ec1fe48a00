"""Build the gfx950 C-ABI library ``tenet/libtenet_amd.so`` in-tree with hipcc.

Usage: ``python tenet-py_amd/build.py``.  Every call recompiles (the build check of this
repository compiles each time by design).
"""

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "tenet" / "libtenet_amd.so"
SOURCES = ["recouple.hip", "gemm.hip", "blas1.hip", "svd.hip", "krylov.hip", "batch.hip", "plan.cpp"]
ARCH = os.environ.get("TENET_OFFLOAD_ARCH", "gfx950")


def _run(cmd, verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature; harmless
    noise = "is not a recognized feature for this target"
    err = "\n".join(line for line in res.stderr.splitlines() if noise not in line)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if err.strip() and verbose:
        print(err, file=sys.stderr)


def build(verbose: bool = False) -> Path:
    """Compile every source to an object (in parallel; the sources share no device code)
    and link them into the shared library."""
    from concurrent.futures import ThreadPoolExecutor
    import tempfile

    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT.with_suffix(".so.tmp")
    flags = [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", str(HERE.parent / "include"),
    ]
    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with tempfile.TemporaryDirectory(prefix="tenet_build_") as d:
        objs = [str(Path(d) / (s + ".o")) for s in SOURCES]
        with ThreadPoolExecutor(jobs) as ex:
            for f in [ex.submit(_run, [hipcc, *flags, "-c", str(CSRC / s), "-o", o], verbose)
                      for s, o in zip(SOURCES, objs)]:
                f.result()
        _run([hipcc, f"--offload-arch={ARCH}", "-shared", *objs, "-o", str(tmp)], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
