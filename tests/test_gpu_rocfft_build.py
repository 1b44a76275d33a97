"""The rocFFT kernels the library runs carry no packed FP32 VALU operations.

rocFFT builds its kernels at run time through hiprtc and also takes prebuilt ones from a system
kernel cache.  With default flags they contain packed FP32 operations (v_pk_fma_f32 and kin),
which on MI355X have given wrong results beside matrix-core work: with two serving slots at 384^3
the last c2r batch of one slot's shell maps changed while the other slot's triangle GEMM ran.
Importing pbi4_amd (or loading libpbi4hip.so) makes rocFFT build every kernel with packed FP32 off
into the library's own cache and skip the system cache; this test runs every kind of transform
the library plans, then disassembles each kernel in that cache (llvm-objdump of code built on
this machine) and checks that none uses a packed FP32 instruction."""
import os
import re
import sqlite3
import subprocess
import tempfile

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu
OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"


def test_rocfft_kernels_built_without_packed_fp32():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pbi4_amd.core import Context
    ctx = Context(0)
    rng = np.random.default_rng(5)
    for N, inter in ((48, True), (48, False), (30, True)):
        pos = torch.from_numpy(rng.random((3, 20000)).astype(np.float32)).cuda()
        d = ctx.density(ctx.assign(pos, N, 3, inter), N, 3, inter, 1000.0, 20000)
        from pbi4_amd.pipeline import itot_count
        itot = itot_count(5, 1.0, 1.0, True)
        ntr = ctx.count_triangles(5, N, 1.0, 1.0, True, False, itot)
        counts = torch.from_numpy((ntr.astype(np.float64) * N ** 3).astype(np.float32)).cuda()
        for mg in (0, N):          # band-limited 2D c2r and full-grid 3D c2r maps
            desc = ctx.bk_desc(5, N, 1.0, 1.0, True, np.float32(2 * np.pi / 1000), itot, 20000,
                               True, (0.0, 0.0, 1.0), 1, mg)
            ctx.bk(d, counts, desc)
    torch.cuda.synchronize()
    path = os.environ["ROCFFT_RTC_CACHE_PATH"]
    assert path == os.environ["PBI4_ROCFFT_CACHE"] and os.path.exists(path), path
    rows = sqlite3.connect(path).execute("select kernel_name, code from cache_v1").fetchall()
    assert len(rows) >= 4
    bad = []
    with tempfile.TemporaryDirectory() as tmp:
        for k, (name, code) in enumerate(rows):
            fn = os.path.join(tmp, f"k{k}.co")
            with open(fn, "wb") as f:
                f.write(code)
            dis = subprocess.run([OBJDUMP, "-d", fn], capture_output=True, text=True, check=True).stdout
            assert "v_" in dis, name
            if re.search(r"\bv_pk_(fma|mul|add)_f32\b", dis):
                bad.append(name)
    assert not bad, bad
