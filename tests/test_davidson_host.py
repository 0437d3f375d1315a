"""Host-side pieces of the block Davidson (CPU): the split-K Gram product
and the zeroth-order diagonal used by the preconditioner."""
import numpy as np
import torch

import elvibrot_amd as E
from elvibrot_amd import psi_io
from elvibrot_amd.davidson import _gram, ho_zeroth_order_diagonal


def test_gram_equals_plain_product_on_row_slices():
    g = torch.Generator().manual_seed(1)
    buf = torch.randn(30, 5003, dtype=torch.float64, generator=g)
    W = torch.randn(6, 5003, dtype=torch.float64, generator=g)
    for A, B in ((buf[:17], buf[:9]), (W, buf[:30]), (W, W), (W[:, :100], buf[:4, :100])):
        ref = A @ B.T
        assert torch.max(torch.abs(_gram(A, B) - ref)).item() <= 1e-12 * torch.max(torch.abs(ref)).item()


def test_zeroth_order_diagonal_is_sum_of_quanta():
    """E0(ib) = sum_k (i_k + 1/2) over the packed row's mode indices
    (get_Ene0_AT_ib_FROM_Basis for unit-frequency HO bases)."""
    b = E.SG4Basis(4, 5)
    e0 = ho_zeroth_order_diagonal(b, "cpu").numpy()
    idx = psi_io.packed_indices(b)
    assert np.array_equal(e0, idx.sum(1) + 0.5 * b.D)


class _DenseOp:
    """The oracle's 6D L=3 Henon-Heiles operator as a dense CPU matrix, with
    the interface the Davidson driver uses; prec != 0 perturbs H.psi by a
    relative 1e-7 (the role of a reduced-precision apply)."""

    def __init__(self, max_nvec):
        from oracle import oracle as O
        self.basis = E.SG4Basis(6, 3)
        o = O.SG4(6, 3)
        H = o.hpsi(np.eye(o.nb_ba), o.grid_hh(0.111803), -0.5 * np.ones(6))
        self.H = torch.tensor(0.5 * (H + H.T))
        self.nb_ba, self.max_nvec = o.nb_ba, max_nvec
        self.g = torch.Generator().manual_seed(3)

    def apply(self, X, Y, prec):
        Y.copy_(X @ self.H)
        if prec:
            Y.add_(1e-7 * torch.randn(Y.shape, dtype=torch.float64, generator=self.g) * Y.abs())


def test_davidson_driver_on_a_dense_operator_with_restarts():
    """The block Davidson (incremental projected matrix, Ritz vectors kept for
    restarts, reduced-precision phase then FP64) on the CPU against the
    eigenvalues of the dense matrix, with a subspace small enough to restart."""
    from elvibrot_amd.davidson import sub_propagation_Davidson
    op = _DenseOp(4)
    ref = np.linalg.eigvalsh(op.H.numpy())[:10]
    for prec in (0, 2):
        r = sub_propagation_Davidson(op, 10, max_it=300, max_WP=22, conv_resi=1e-7,
                                     conv_ene=1e-10, prec=prec, device="cpu")
        assert r.converged.all()
        assert np.max(np.abs(r.E - ref)) < 1e-9, (prec, r.E - ref)
        assert max(h[1] for h in r.history) <= 22
        X = r.X.numpy()
        assert np.max(np.abs(X @ X.T - np.eye(10))) < 1e-10
