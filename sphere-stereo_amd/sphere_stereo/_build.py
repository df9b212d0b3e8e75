"""Builds libsphere_stereo.so for gfx950 with hipcc (no CMake, no JIT cache).

The library is written next to this file so that it travels with the source
tree.  Run as ``python sphere-stereo_amd/sphere_stereo/_build.py`` or through
``__graft_entry__.build()``; both load this file by path, because importing
the package loads the library (and fails on a stale one).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
LIB = os.path.join(HERE, "libsphere_stereo.so")
SOURCES = ["api.hip", "frame.hip", "ragged.hip", "isb.hip", "sweep.hip", "stitch.hip", "eval.hip", "input.hip", "grad.hip"]

# -ffp-contract=off and correctly rounded fp32 division/sqrt keep every
# operation a single IEEE round-to-nearest step (bit-exact with the oracle);
# packed fp32 instructions are disabled (they must not be generated).
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    root = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = os.path.join(root, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def build(verbose: bool = False, extra_flags=None) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + list(extra_flags or []) + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
