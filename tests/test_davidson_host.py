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
