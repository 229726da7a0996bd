"""Build the oracle's C restatement (oracle/_build/libnbody_ref.so) with gcc.

TEST INFRASTRUCTURE ONLY (CPU baseline and second oracle).  The flags mirror
numba's @njit(parallel=True, fastmath=True): -O3 -ffast-math -fopenmp, and
-march=x86-64-v3 (AVX2/FMA) so the library runs on any current x86 host.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "libnbody_ref.so")


def build(verbose=True):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["gcc", "-O3", "-ffast-math", "-fopenmp", "-march=x86-64-v3", "-fPIC", "-shared",
           os.path.join(HERE, "nbody_ref.c"), "-o", OUT + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print(f"[oracle build] {OUT}")
    return OUT


if __name__ == "__main__":
    build()
