"""Compile libmplicits_hip.so (HIP, gfx950) and the C restatements of the oracle.

Every .hip file under csrc/ is compiled to an object with hipcc for gfx950 and
linked into mplicits/libmplicits_hip.so, next to the Python package that loads
it with ctypes.  Run directly (``python m-plicits_amd/build_lib.py``) or through
``__graft_entry__.build()``.
"""
import concurrent.futures as cf
import os
import os.path as osp
import subprocess
import sys

HERE = osp.dirname(osp.abspath(__file__))
CSRC = osp.join(HERE, "csrc")
OBJ = osp.join(HERE, "_build")
LIB = osp.join(HERE, "mplicits", "libmplicits_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: keep the compiler from forming v_pk_{add,mul,fma}_f32.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


# Per-file additions.  level.hip, trace.hip, batch.hip: let the MFMAs accumulate in VGPRs (the
# epilogue then reads the accumulators without v_accvgpr_read copies).
FILE_FLAGS = {
    "level.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "trace.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "render.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "batch.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}


def _compile(src):
    obj = osp.join(OBJ, osp.basename(src) + ".o")
    deps = [src] + [osp.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    deps.append(osp.join(HERE, "..", "include", "mplicits.h"))
    deps.append(osp.abspath(__file__))  # flags live here
    if osp.exists(obj) and all(osp.getmtime(obj) >= osp.getmtime(d) for d in deps):
        return obj, ""
    cmd = [HIPCC] + HIP_FLAGS + FILE_FLAGS.get(osp.basename(src), []) + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{p.stdout}\n{p.stderr}")
    return obj, p.stderr


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(osp.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for o, log in results:
            noisy = [l for l in log.splitlines() if "packed-fp32-ops" not in l]
            if noisy:
                print(o, "\n".join(noisy))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stdout}\n{p.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
