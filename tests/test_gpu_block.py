"""The dense block kernels of the device Davidson (sg4_block_gram,
sg4_block_combine) through the C ABI against numpy products: row blocks cut
from larger buffers (row strides above n), an inner dimension that is not a
multiple of the 32-column tile, the largest shapes the kernels take, the
fixed-order sum (two grams bitwise equal) and the refusals."""
import ctypes as C

import numpy as np
import pytest

import elvibrot_amd as E

pytestmark = pytest.mark.gpu

N = 100_003


def _plan():
    return E.SG4Operator(E.SG4Basis(3, 2), max_nvec=4)


def _gram(op, A, B, work=None):
    import torch
    m, q = A.shape[0], B.shape[0]
    need = C.c_int64()
    E._capi.check(E.lib().sg4_block_gram_work(m, q, C.byref(need)))
    work = torch.empty(need.value, device=A.device, dtype=torch.float64) if work is None else work
    out = torch.empty(m, q, device=A.device, dtype=torch.float64)
    rc = E.lib().sg4_block_gram(op._h, m, q, A.shape[1], C.c_void_p(A.data_ptr()), A.stride(0),
                                C.c_void_p(B.data_ptr()), B.stride(0), C.c_void_p(out.data_ptr()),
                                C.c_void_p(work.data_ptr()), work.numel())
    return rc, out


@pytest.mark.parametrize("m,q", [(1, 1), (3, 84), (16, 84), (84, 84), (124, 4), (7, 13), (20, 36)])
def test_gram_against_numpy(cuda, m, q):
    import torch
    op = _plan()
    g = torch.Generator(device=cuda).manual_seed(m * 1000 + q)
    buf = torch.randn(m + q + 2, N + 37, device=cuda, dtype=torch.float64, generator=g)
    A, B = buf[:m, 5:5 + N], buf[m + 1:m + 1 + q, 11:11 + N]
    rc, got = _gram(op, A, B)
    assert rc == 0
    ref = A.cpu().numpy() @ B.cpu().numpy().T
    err = np.max(np.abs(got.cpu().numpy() - ref)) / np.max(np.abs(ref))
    print(f"gram {m}x{q}: {err:.2e}")
    assert err < 1e-13
    rc2, again = _gram(op, A, B)          # a fresh work array of the same size
    assert rc2 == 0 and torch.equal(got, again)
    # a work array for fewer workgroups gives the same sums to rounding
    rc3, few = _gram(op, A, B, work=torch.empty(3 * m * q, device=cuda, dtype=torch.float64))
    assert rc3 == 0 and np.max(np.abs(few.cpu().numpy() - ref)) / np.max(np.abs(ref)) < 1e-13


@pytest.mark.parametrize("k,d,beta", [(1, 1, 0.0), (36, 84, 0.0), (20, 48, 1.0), (5, 7, -0.5),
                                      (17, 3, 0.0), (128, 64, 1.0)])
def test_combine_against_numpy(cuda, k, d, beta):
    import torch
    op = _plan()
    g = torch.Generator(device=cuda).manual_seed(k * 100 + d)
    X = torch.randn(d, N + 9, device=cuda, dtype=torch.float64, generator=g)[:, 4:4 + N]
    U = torch.randn(k, d, device=cuda, dtype=torch.float64, generator=g)
    Ybuf = torch.randn(k, N + 3, device=cuda, dtype=torch.float64, generator=g)
    Y = Ybuf[:, 1:1 + N]
    if beta == 0.0:
        Y.fill_(float("nan"))          # not read with beta = 0
    Y0 = Y.cpu().numpy().copy()
    rc = E.lib().sg4_block_combine(op._h, k, d, N, C.c_void_p(U.data_ptr()), C.c_void_p(X.data_ptr()),
                                   X.stride(0), beta, C.c_void_p(Y.data_ptr()), Y.stride(0))
    assert rc == 0
    ref = U.cpu().numpy() @ X.cpu().numpy() + (beta * Y0 if beta != 0.0 else 0.0)
    err = np.max(np.abs(Y.cpu().numpy() - ref)) / np.max(np.abs(ref))
    print(f"combine {k}x{d} beta {beta}: {err:.2e}")
    assert err < 1e-13
    assert torch.equal(Ybuf[:, 0].isnan(), torch.zeros(k, dtype=torch.bool, device=cuda))


def test_block_kernels_refuse_bad_input(cuda):
    import torch
    op = _plan()
    need = C.c_int64()
    assert E.lib().sg4_block_gram_work(125, 4, C.byref(need)) == -1
    assert E.lib().sg4_block_gram_work(100, 100, C.byref(need)) == -1      # 25 x 25 > 512 blocks
    A = torch.zeros(4, 1000, device=cuda, dtype=torch.float64)
    out = torch.zeros(4, 4, device=cuda, dtype=torch.float64)
    host = np.zeros(4 * 1000)
    rc = E.lib().sg4_block_gram(op._h, 4, 4, 1000, host.ctypes.data_as(C.c_void_p), 1000,
                                C.c_void_p(A.data_ptr()), 1000, C.c_void_p(out.data_ptr()),
                                C.c_void_p(out.data_ptr()), 16)
    assert rc == -1
    U = torch.zeros(100, 100, device=cuda, dtype=torch.float64)
    X = torch.zeros(100, 1000, device=cuda, dtype=torch.float64)
    assert E.lib().sg4_block_combine(op._h, 100, 100, 1000, C.c_void_p(U.data_ptr()),
                                     C.c_void_p(X.data_ptr()), 1000, 0.0, C.c_void_p(X.data_ptr()),
                                     1000) == -1                                  # k d > 8192
    assert E.lib().sg4_block_combine(op._h, 4, 4, 1000, C.c_void_p(U.data_ptr()),
                                     C.c_void_p(X.data_ptr()), 1000, 0.0,
                                     C.c_void_p(X[2:].data_ptr()), 1000) == -1   # Y overlaps X
    assert "overlap" in E.lib().sg4_last_error().decode()
