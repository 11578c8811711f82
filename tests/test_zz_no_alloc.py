"""mlv_reduce_noalloc (the reduction that never allocates device memory:
its partials live in the library's fixed device array) against mlv_reduce
with a plan-less call, through the C ABI: bit-identical results for every
operation and both precisions, on two calls in a row with the second input
smaller than the first (partials left by the first call must not leak into
the second)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(L, name, *args):
    L.call(name, *args)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_reduce_noalloc_equals_reduce(mlv, dtype):
    import torch

    L = mlv._lib
    tdt = torch.float64 if dtype == "f64" else torch.float32
    code = L.F64 if dtype == "f64" else L.F32
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    stream = L.stream_handle()
    for n in (3_000_017, 1_001):   # all first-stage blocks, then only four
        a = torch.randn(n, generator=gen, device="cuda", dtype=tdt)
        b = torch.randn(n, generator=gen, device="cuda", dtype=tdt)
        a[n // 3] = 7.5   # a known maximum
        ah = a.double().cpu().numpy()
        bh = b.double().cpu().numpy()
        for op, scale in ((L.RED_MAX, 1.0), (L.RED_SUM, 1.0), (L.RED_SUM, 1.0 / n),
                          (L.RED_SUMSQ2, 0.5), (L.RED_DOT, 1.0)):
            print(f"[noalloc] {dtype} n={n} op={op}", flush=True)
            ref = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
            got = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
            bp = ctypes.c_void_p(b.data_ptr())
            _call(L, "mlv_reduce", None, op, code, ctypes.c_void_p(a.data_ptr()), bp, n,
                  float(scale), ctypes.c_void_p(ref.data_ptr()), stream)
            _call(L, "mlv_reduce_noalloc", op, code, ctypes.c_void_p(a.data_ptr()), bp, n,
                  float(scale), ctypes.c_void_p(got.data_ptr()), stream)
            r, g = ref.item(), got.item()
            assert g == r, (op, g, r)
            # and the value itself, against NumPy in FP64
            want = {L.RED_MAX: ah.max(), L.RED_SUM: ah.sum() * scale,
                    L.RED_SUMSQ2: (ah * ah + bh * bh).sum() * scale,
                    L.RED_DOT: (ah * bh).sum()}[op]
            assert abs(g - want) <= 1e-12 * max(1.0, abs(want)) * (1 if op == L.RED_MAX else 10)


def test_reduce_noalloc_rejects_bad_arguments(mlv):
    L = mlv._lib
    f = L.lib().mlv_reduce_noalloc
    one = ctypes.c_void_p(1)   # never dereferenced: every call fails validation
    assert f(L.RED_SUM, L.F64, None, None, 4, 1.0, one, None) == L.MLV_ERR_ARG
    assert f(L.RED_SUM, L.F64, one, None, 0, 1.0, one, None) == L.MLV_ERR_ARG
    assert f(L.RED_DOT, L.F64, one, None, 4, 1.0, one, None) == L.MLV_ERR_ARG
    assert f(L.RED_SUM, 7, one, None, 4, 1.0, one, None) == L.MLV_ERR_ARG
    assert f(9, L.F64, one, None, 4, 1.0, one, None) == L.MLV_ERR_ARG
