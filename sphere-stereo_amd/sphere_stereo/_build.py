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


def build(verbose: bool = False, extra_flags=None, jobs=None) -> str:
    """Compiles every source to an object in parallel (up to `jobs` hipcc
    processes, default: the CPU count, at most 16), then links the library."""
    import concurrent.futures
    import tempfile

    flags = [f for f in HIPCC_FLAGS if f != "-shared"] + list(extra_flags or [])
    jobs = jobs or min(16, os.cpu_count() or 1)
    tmp = LIB + ".tmp"
    with tempfile.TemporaryDirectory(prefix="ss_build_") as obj_dir:
        def compile_one(src):
            obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
            cmd = [hipcc()] + flags + ["-c", "-o", obj, os.path.join(CSRC, src)]
            if verbose:
                print(" ".join(cmd), flush=True)
            res = subprocess.run(cmd, capture_output=True, text=True)
            if res.returncode != 0:
                raise RuntimeError("hipcc failed on %s:\n%s%s" % (src, res.stdout, res.stderr))
            return obj

        # the slowest sources first, so that they start at once
        order = sorted(SOURCES, key=lambda f: -os.path.getsize(os.path.join(CSRC, f)))
        with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
            objs = dict(zip(order, ex.map(compile_one, order)))
        cmd = [hipcc()] + HIPCC_FLAGS + list(extra_flags or []) + ["-o", tmp] + [objs[s] for s in SOURCES]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("hipcc link failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


PROBE_SRC = os.path.join(CSRC, "probes", "issue_rates.hip")
PROBE_BIN = os.path.join(HERE, "issue_rates")


def build_probes(verbose: bool = False) -> str:
    """Compiles the issue-rate probe (csrc/probes/issue_rates.hip), a
    stand-alone gfx950 program that `bench.py --issue-rates` runs."""
    flags = [f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    tmp = PROBE_BIN + ".tmp"
    cmd = [hipcc()] + flags + ["-o", tmp, PROBE_SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s%s" % (PROBE_SRC, res.stdout, res.stderr))
    os.replace(tmp, PROBE_BIN)
    return PROBE_BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_probes(verbose="-v" in sys.argv))
