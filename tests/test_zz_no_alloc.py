"""The allocation-free entry point sg4_apply_noalloc (staging blocks in the
library's device scratch array) against sg4_apply, through the C ABI: two
calls in a row, the second with a smaller block, in FP64 (1e-12 relative)
and in the split-precision mode (same kernels as sg4_apply: 1e-12); an
unprepared block size and an oversized block are refused."""
import ctypes as C

import numpy as np
import pytest

import elvibrot_amd as E

pytestmark = pytest.mark.gpu


def call_noalloc(op, x, y, prec):
    return E.lib().sg4_apply_noalloc(op._h, x.shape[0], C.c_void_p(x.data_ptr()),
                                     C.c_void_p(y.data_ptr()), prec)


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_F32X2])
def test_noalloc_equals_apply(cuda, prec):
    import torch
    print(f"noalloc prec {prec}", flush=True)
    D, L = 6, 5
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=16)
    g = torch.Generator(device=cuda).manual_seed(11)
    X16 = torch.randn(16, b.nb_ba, device=cuda, dtype=torch.float64, generator=g)
    X5 = torch.randn(5, b.nb_ba, device=cuda, dtype=torch.float64, generator=g)
    R16, R5 = torch.empty_like(X16), torch.empty_like(X5)
    op.apply(X16, R16, prec)          # the existing entry point (prepares the plan)
    op.apply(X5, R5, prec)
    Y16, Y5 = torch.empty_like(X16), torch.empty_like(X5)
    E._capi.check(call_noalloc(op, X16, Y16, prec))
    E._capi.check(call_noalloc(op, X5, Y5, prec))
    torch.cuda.synchronize()
    rel = lambda a, r: (torch.max(torch.abs(a - r)) / torch.max(torch.abs(r))).item()
    e16, e5 = rel(Y16, R16), rel(Y5, R5)
    print(f"noalloc vs apply: {e16:.2e} (16 vectors), {e5:.2e} (5 vectors)")
    assert e16 < 1e-12 and e5 < 1e-12
    # a block whose padded width (3 -> 4 vectors) the plan was never prepared
    # for is refused, not built (16 and 5 -> 8 were prepared above)
    X3 = torch.randn(3, b.nb_ba, device=cuda, dtype=torch.float64, generator=g)
    assert call_noalloc(op, X3, torch.empty_like(X3), prec) == -5   # SG4_ERR_STATE


def test_noalloc_refuses_block_beyond_scratch(cuda):
    import torch
    b = E.SG4Basis(8, 8)                       # nb_ba = 265729; x 64 > 2^24
    op = E.SG4Operator(b, max_nvec=64)
    X = torch.zeros(64, b.nb_ba, device=cuda, dtype=torch.float64)
    assert call_noalloc(op, X, torch.empty_like(X), E.PREC_FP64) == -1   # SG4_ERR_ARG
    assert "exceeds" in E.lib().sg4_last_error().decode()
