"""Build liblakde_amd.so for gfx950 with hipcc (no cmake/ninja needed).

    python -m lakde_amd.build [-j JOBS]

Sources live in lakde_amd/csrc.  The dimension-templated fused kernels are
explicitly instantiated in generated translation units (one per group of
dimensions, under lakde_amd/_build/) so hipcc compiles them in parallel.
"""

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblakde_amd.so")
ARCH = os.environ.get("LAKDE_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

CFLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    # exact, context-independent rounding (the E-step ballot and compaction
    # recompute the same values) and no packed-FP32 instruction selection
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-command-line-argument",
]

SOURCES = ["ops_ref.hip", "posterior.hip", "post_small.hip", "fused.hip", "bf16.hip",
           "backward.hip", "step.hip", "knn.hip"]
D_GROUPS = [(1, 6), (7, 12), (13, 18), (19, 24), (25, 32), (33, 40), (41, 48), (49, 56),
            (57, 63)]


def _gen_inst(lo, hi):
    path = os.path.join(BUILD, f"fused_d{lo}_{hi}.hip")
    lines = ['#include "%s"' % os.path.join(CSRC, "fused_impl.hpp"),
             '#include "%s"' % os.path.join(CSRC, "screen.hpp"),
             '#include "%s"' % os.path.join(CSRC, "mfma32.hpp"), "namespace lakde {"]
    for d in range(lo, hi + 1):
        lines += [
            f"template int launch_estep_main<{d}>(const EParams &, int, hipStream_t);",
            f"template int launch_estep_compact<{d}>(const EParams &, hipStream_t);",
            f"template int launch_pred_main<{d}>(const PParams &, hipStream_t);",
            f"template int launch_bwd_y<{d}>(const PParams &, const float *, const float *, "
            "float *, hipStream_t);",
            f"template int launch_estep_screen<{d}>(const EParams &, const SParams &, int, "
            "hipStream_t);",
            f"template int launch_pack_screen<{d}>(const float *, int64_t, const float *, "
            "uint8_t *, hipStream_t);",
            f"template int launch_screen_points<{d}>(const float *, int64_t, const float *, "
            "__bf16 *, float *, hipStream_t);",
            f"template int launch_screen_knn_bound<{d}>(const float *, const float *, int64_t, "
            "const void *, int, int64_t, int, float *, hipStream_t);",
            f"template int launch_pred_screen<{d}>(const PParams &, const uint8_t *, "
            "const __bf16 *, const float *, float, hipStream_t);",
            f"template int launch_estep_m32<{d}>(const EParams &, const float *, const float *, "
            "int, hipStream_t);",
            f"template int launch_pack_m32<{d}>(const float *, int64_t, const float *, float *, "
            "hipStream_t);",
        ]
    lines.append("}")
    text = "\n".join(lines) + "\n"
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)
    return path


# per-source flags: the matrix-core kernels keep their MFMA accumulators in
# ordinary VGPRs (the epilogue's squares read them without v_accvgpr_read);
# the dimension-instantiated units hold the screening E-step's MFMA kernels
MFMA_VGPR = ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]
EXTRA_FLAGS = {"bf16.hip": MFMA_VGPR}


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
    name = os.path.basename(src)
    extra = EXTRA_FLAGS.get(name, MFMA_VGPR if name.startswith("fused_d") else [])
    cmd = [HIPCC] + CFLAGS + extra + ["-c", src, "-o", obj]
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
    return obj, time.time() - t


def build(jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    srcs += [_gen_inst(lo, hi) for lo, hi in D_GROUPS]
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    objs = []
    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, dt in ex.map(_compile, srcs):
            objs.append(obj)
            if verbose:
                print(f"  compiled {os.path.basename(obj)} in {dt:.1f}s", flush=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s", flush=True)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    args = ap.parse_args()
    try:
        build(args.j)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
