"""Build recipe of libreproblas_hip.so (the C-ABI library with the gfx950 kernels).

Every source is compiled by hipcc for gfx950 only, with FMA contraction off
(binned arithmetic must never fuse a*b+c) and packed-FP32 instruction
selection disabled; objects go to reproblas_amd/_build/, the library to
reproblas_amd/libreproblas_hip.so (in-tree, so it travels with the repo).

    python -m reproblas_amd.build        # or reproblas_amd.build.build()
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libreproblas_hip.so")

# A unit is a source file, or (source file, preprocessor definitions, object
# name) for a file compiled more than once.
SOURCES = [
    "runtime.hip",
    "host_binned.hip",
    "binned_vec.hip",
    "blas1_d.hip",
    "blas1_f.hip",
    "blas1_seg.hip",
    *[("blas1_inst.hip", {"RB_INST_TERM": str(t), "RB_INST_TYPE": ty}, f"blas1_{name}_{ty[0]}")
      for t, name in enumerate(("sum", "asum", "dot", "ssq")) for ty in ("double", "float")],
    "complex_d.hip",
    "complex_f.hip",
    *[("complex_inst.hip", {"RB_INST_TERM": str(t), "RB_INST_TYPE": ty}, f"complex_{name}_{ty[0]}")
      for t, name in enumerate(("sum", "dotu", "dotc", "asum", "ssq")) for ty in ("double", "float")],
    "complex_mat.hip",
    "amax.hip",
    "gemv.hip",
    "gemm.hip",
    "batch.hip",
    "api.hip",
    "host_complex.hip",
    "api_complex.hip",
    "noalloc.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def _unit(u):
    """(source, definitions, object path) of a SOURCES entry."""
    if isinstance(u, str):
        return u, {}, os.path.join(OBJDIR, os.path.splitext(u)[0] + ".o")
    src, defs, name = u
    return src, defs, os.path.join(OBJDIR, name + ".o")


def _stale(src, obj):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    deps += [os.path.join(HERE, "..", "include", h) for h in os.listdir(os.path.join(HERE, "..", "include"))]
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(unit, incremental=False):
    src, defs, obj = _unit(unit)
    if incremental and not _stale(src, obj):
        return obj, ""
    cmd = [HIPCC, *FLAGS, *[f"-D{k}={v}" for k, v in defs.items()], "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(jobs=None, verbose=False, incremental=False):
    """Compile every HIP source for gfx950 and link the shared library.

    incremental=True skips sources whose object is newer than the source and
    every header (for development); the default recompiles everything."""
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    units = [u for u in SOURCES if os.path.exists(os.path.join(CSRC, _unit(u)[0]))]
    # longest units first, so the pool finishes together
    heavy = ("blas1_", "complex_", "gemm", "gemv")
    units.sort(key=lambda u: not any(h in _unit(u)[2] for h in heavy))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda f: _compile(f, incremental), units))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
