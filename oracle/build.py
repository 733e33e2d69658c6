"""Builds the oracle (oracle/liboracle.so) from oracle/lens.c with the reference's floating-point
rules (-ffp-contract=off -fno-math-errno -fno-trapping-math, Artful CMakeLists.txt:54-58); and,
where the reference tree is present, oracle/_ref/libartful_ref.so from it (oracle/ref/Makefile), and
oracle/_ref/libartful_routeb.so (the reference's tree with the HIP backend added, route B).
Test infrastructure only."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
OUT = HERE / "liboracle.so"
REF = Path(os.environ.get("ARTF_REFERENCE", "/root/reference"))


def build_oracle() -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["gcc", "-std=c11", "-O2", "-fPIC", "-shared", "-Wall", "-ffp-contract=off", "-fno-math-errno",
           "-fno-trapping-math", f"-I{ROOT / 'include'}", str(HERE / "lens.c"), "-o", str(tmp), "-lm", "-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


def build_reference() -> Path | None:
    """The reference's own code, for golden vectors and the CPU baseline; None where it is absent."""
    if not (REF / "src" / "software" / "lens.inc").exists():
        return None
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-s", "-C", str(HERE / "ref"), f"-j{jobs}", f"REF={REF}"], check=True)
    # route B: the reference's tree with this repository's HIP backend added (needs the kernels' objects,
    # which artful_amd/build.py leaves in _build/artful_hip)
    subprocess.run(["make", "-s", "-C", str(HERE / "ref"), f"-j{jobs}", f"REF={REF}", "routeb"], check=True)
    return HERE / "_ref" / "libartful_ref.so"


if __name__ == "__main__":
    print(build_oracle())
    if "--no-ref" not in sys.argv:
        print(build_reference())
