"""pbi4_triangle_sums_noalloc (split-K partials in the library's static device array) against
pbi4_triangle_sums with a caller workspace, through the C ABI: the same plan and kernels, so the
sums must be bitwise equal, on two calls in a row (the second smaller, so it runs on what the
first left in the array), plus the refusal of a shape whose partials exceed the array."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def test_noalloc_triangle_sums_matches_workspace_entry_point():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pbi4_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.pbi4_ctx_create(0, ctypes.byref(h)) == 0
    try:
        g = torch.Generator(device="cuda")
        g.manual_seed(11)
        # (shells, multipoles, K, precision): 16 shells x 3 on 100^3 (the serving shape), then a
        # smaller FP32 problem, then a smaller split-BF16 one
        for nb, nell, K, prec in ((16, 3, 100 ** 3, _lib.PREC_BF16X3), (6, 1, 20000, _lib.PREC_FP32),
                                  (5, 3, 4096, _lib.PREC_BF16X3)):
            Ks = (K + 63) // 64 * 64
            maps = torch.randn((nb * nell, Ks), generator=g, device="cuda", dtype=torch.float32)
            maps[:, K:] = 0
            npair, ncol = nb * (nb + 1) // 2, nb * nell
            ws_bytes = L.pbi4_triangle_sums_workspace(nb, nell, K, Ks, prec)
            assert 0 < ws_bytes - 256 <= _lib.NOALLOC_SCRATCH_BYTES
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
            c_ws = torch.empty((npair, ncol), dtype=torch.float64, device="cuda")
            c_na = torch.full((npair, ncol), float("nan"), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            print(f"[noalloc] nb={nb} nell={nell} K={K} prec={prec}", flush=True)
            assert L.pbi4_triangle_sums(h, nb, nell, K, Ks, prec, _p(maps), _p(c_ws), _p(ws),
                                        ws_bytes, None) == 0
            assert L.pbi4_triangle_sums_noalloc(h, nb, nell, K, Ks, prec, _p(maps), _p(c_na),
                                                None) == 0, L.pbi4_last_error(h)
            torch.cuda.synchronize()
            assert torch.equal(c_ws, c_na)
            # and both are the triple sums (FP64 check on a few pairs)
            m = maps[:, :K].double().cpu().numpy()
            for p, (i, j) in ((0, (0, 0)), (npair - 1, (nb - 1, nb - 1))):
                exp = m @ (m[i] * m[j])
                tol = (3e-5 if prec == _lib.PREC_BF16X3 else 2e-6) * 5 * np.sqrt(
                    np.sum((m[i] * m[j])[None, :] ** 2 * m ** 2, axis=1))
                assert np.all(np.abs(c_na[p].cpu().numpy() - exp) <= tol + 1e-12)
        # 48 shells x 3 on 150^3 need ~300 MB of partials: refused, nothing launched
        K = 150 ** 3
        Ks = (K + 63) // 64 * 64
        need = L.pbi4_triangle_sums_workspace(48, 3, K, Ks, _lib.PREC_BF16X3) - 256
        assert need > _lib.NOALLOC_SCRATCH_BYTES
        maps = torch.zeros((48 * 3, Ks), device="cuda", dtype=torch.float32)
        c = torch.zeros((48 * 49 // 2, 48 * 3), device="cuda", dtype=torch.float64)
        rc = L.pbi4_triangle_sums_noalloc(h, 48, 3, K, Ks, _lib.PREC_BF16X3, _p(maps), _p(c), None)
        assert rc == -4
        assert b"scratch" in L.pbi4_last_error(h)
    finally:
        L.pbi4_ctx_destroy(h)
