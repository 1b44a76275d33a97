"""Builds libpbi4hip.so in-tree with hipcc for gfx950 (no JIT, no torch extension machinery).

    python -m pbi4_amd.build            # or pbi4_amd.build.build()

Objects go to pbi4_amd/_build/, the library to pbi4_amd/libpbi4hip.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpbi4hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
SOURCES = ["assign.hip", "spectrum.hip", "triangle.hip", "capi.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
         # packed FP32 VALU ops are kept out of the device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(HERE, "..", "include")]


def _compile(src):
    obj = os.path.join(HERE, "_build", os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stderr or "").splitlines() if "packed-fp32-ops" not in l]
    return obj, r.returncode, "\n".join(msgs)


def build(verbose=False):
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        results = list(ex.map(_compile, SOURCES))
    failed = False
    for (obj, rc, msg), src in zip(results, SOURCES):
        if msg.strip() and (verbose or rc != 0):
            print(f"[{src}]\n{msg}", file=sys.stderr)
        failed |= rc != 0
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp,
           *[r[0] for r in results], "-L" + os.path.join(ROCM, "lib"), "-lrocfft",
           "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stderr, file=sys.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
