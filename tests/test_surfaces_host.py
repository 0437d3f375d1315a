"""Block bookkeeping of SG4Surfaces on the CPU, with dense stand-ins for
the single-surface plans: every (i, j) block applied to component j and
summed into component i, a block object shared by (i, j) and (j, i) applied
once to both components, blocks wider than a plan's max_nvec chunked, zero
blocks (None) skipped, against the dense block matrix."""
import numpy as np
import torch

from elvibrot_amd.surfaces import SG4Surfaces


class _Basis:
    def __init__(self, nb):
        self.nb_ba = nb


class _Dense:
    def __init__(self, M, basis, max_nvec):
        self.M = torch.tensor(M)
        self.basis, self.max_nvec, self._h = basis, max_nvec, None
        self.calls = []

    def apply(self, X, Y, prec=0):
        assert X.shape[0] <= self.max_nvec and X.is_contiguous()
        self.calls.append(X.shape[0])
        Y.copy_(X @ self.M.T)


def test_block_sums_pairs_and_chunks():
    rng = np.random.default_rng(1)
    nb, b = 7, _Basis(7)
    sym = lambda: (lambda A: A + A.T)(rng.standard_normal((nb, nb)))
    H = [_Dense(sym(), b, 4) for _ in range(3)]
    C01 = _Dense(rng.standard_normal((nb, nb)), b, 8)     # shared by (0,1) and (1,0)
    C12 = _Dense(rng.standard_normal((nb, nb)), b, 3)
    C21 = _Dense(rng.standard_normal((nb, nb)), b, 3)
    blocks = [[H[0], C01, None], [C01, H[1], C12], [None, C21, H[2]]]
    S = SG4Surfaces(blocks)
    assert S.nb_ba == 3 * nb and S.max_nvec == 3
    big = np.zeros((3 * nb, 3 * nb))
    for i in range(3):
        for j in range(3):
            if blocks[i][j] is not None:
                big[i * nb:(i + 1) * nb, j * nb:(j + 1) * nb] = blocks[i][j].M.numpy()
    for nvec in (1, 4, 5):
        for o in H + [C01, C12, C21]:
            o.calls.clear()
        x = torch.tensor(rng.standard_normal((nvec, 3 * nb)))
        y = torch.full_like(x, np.nan)
        S.apply(x, y)
        assert np.allclose(y.numpy(), x.numpy() @ big.T, rtol=0, atol=1e-12)
        # the shared block: one launch of 2 nvec vectors when they fit its 8,
        # else one launch per side
        assert C01.calls == ([2 * nvec] if 2 * nvec <= 8 else [nvec, nvec])
        assert all(c <= 4 for c in H[0].calls) and sum(H[0].calls) == nvec
        assert sum(C12.calls) == nvec and max(C12.calls) <= 3


def test_zeroth_order_diagonal_repeats_per_surface():
    import elvibrot_amd as E
    b = E.SG4Basis(3, 3)
    from elvibrot_amd.davidson import ho_zeroth_order_diagonal
    H = _Dense(np.eye(b.nb_ba), b, 4)
    S = SG4Surfaces([[H, None], [None, H]])
    d = S.zeroth_order_diagonal("cpu")
    d1 = ho_zeroth_order_diagonal(b, "cpu")
    assert torch.equal(d, torch.cat([d1, d1]))
