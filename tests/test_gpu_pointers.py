"""The device-pointer entry points accept ordinary (coarse-grained) device
memory of the plan's GPU only: their kernels add into the outputs with
hardware floating-point atomics, which are correct on that memory alone.
Host, pinned, managed and fine-grained blocks are refused with SG4_ERR_ARG
before anything is launched, and a refused lookup leaves no HIP error behind
for the next call."""
import ctypes as C

import numpy as np
import pytest

import elvibrot_amd as E
from elvibrot_amd import _capi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _hip():
    # the HIP runtime torch has loaded (same soname as the one the library links)
    h = C.CDLL("libamdhip64.so.7")
    h.hipMallocManaged.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    h.hipExtMallocWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _plan(nvec=4):
    D, L = 4, 4
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, nvec, potential=("henon_heiles", 0.111803))
    return b, op


def _raw_apply(op, fn, nvec, psi_ptr, hpsi_ptr):
    return getattr(_capi.lib(), fn)(op._h, nvec, C.c_void_p(psi_ptr), C.c_void_p(hpsi_ptr),
                                   E.PREC_FP64)


def _err():
    return _capi.lib().sg4_last_error().decode()


def test_device_blocks_accepted_interior_pointers_too(cuda):
    import torch
    b, op = _plan()
    o = O.SG4(4, 4)
    psi = np.random.default_rng(1).standard_normal((b.nb_ba, 3))
    ref = o.hpsi(psi, o.grid_hh(0.111803), -0.5 * np.ones(4))
    big = torch.zeros(2 * b.nb_ba * 3 + 17, dtype=torch.float64, device=cuda)
    x = big[5:5 + b.nb_ba * 3].view(b.nb_ba, 3)          # inside a larger allocation
    x.copy_(torch.tensor(psi))
    y = torch.empty_like(x)
    op.apply_rows(x, y)
    torch.cuda.synchronize()
    assert np.max(np.abs(y.cpu().numpy() - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.parametrize("fn", ["sg4_apply", "sg4_apply_rows", "sg4_apply_graph",
                                "sg4_apply_noalloc"])
def test_host_and_pinned_blocks_refused(cuda, fn):
    import torch
    b, op = _plan()
    dev = torch.zeros(b.nb_ba, 2, dtype=torch.float64, device=cuda)
    out = torch.empty_like(dev)
    pinned = torch.zeros(b.nb_ba, 2, dtype=torch.float64, pin_memory=True)
    plain = np.zeros((b.nb_ba, 2))
    print(f"entry {fn}", flush=True)
    assert _raw_apply(op, fn, 2, pinned.data_ptr(), out.data_ptr()) == _capi.ERR_ARG
    assert "psi is host" in _err()
    assert _raw_apply(op, fn, 2, dev.data_ptr(), pinned.data_ptr()) == _capi.ERR_ARG
    assert "hpsi is host" in _err()
    # ordinary host memory: the runtime reports it as unregistered host
    # memory (or does not know it); refused either way
    assert _raw_apply(op, fn, 2, plain.ctypes.data, out.data_ptr()) == _capi.ERR_ARG
    assert "psi is host" in _err() or "psi is not memory allocated by HIP" in _err()
    # nothing sticky: an ordinary call right after works
    op.apply_rows(dev, out)
    torch.cuda.synchronize()
    assert torch.all(out == 0)


def test_managed_and_fine_grained_blocks_refused(cuda):
    import torch
    b, op = _plan()
    h = _hip()
    nbytes = 8 * b.nb_ba * 2
    out = torch.empty(b.nb_ba, 2, dtype=torch.float64, device=cuda)
    managed, fine = C.c_void_p(), C.c_void_p()
    assert h.hipMallocManaged(C.byref(managed), nbytes, 1) == 0      # hipMemAttachGlobal
    assert h.hipExtMallocWithFlags(C.byref(fine), nbytes, 1) == 0    # hipDeviceMallocFinegrained
    try:
        assert _raw_apply(op, "sg4_apply_rows", 2, managed.value, out.data_ptr()) == _capi.ERR_ARG
        assert "managed" in _err()
        assert _raw_apply(op, "sg4_apply_rows", 2, out.data_ptr(), fine.value) == _capi.ERR_ARG
        assert "fine-grained" in _err()
        mu = torch.zeros(op.basis.sum_nq, dtype=torch.float64, device=cuda)
        assert _capi.lib().sg4_plan_set_field(op._h, C.c_void_p(fine.value), 0.5) == _capi.ERR_ARG
        op.set_field(mu, 0.0)
        op.set_field(None, 0.0)
    finally:
        torch.cuda.synchronize()
        assert h.hipFree(managed) == 0 and h.hipFree(fine) == 0
