"""Operators over several electronic surfaces (nb0 > 1) from single-surface
SG4 plans.

The reference's type-1 action of a Smolyak term on nb0 components
(sub_TabOpPsi_OF_ONEDP_FOR_SGtype4, SRC/sub_Operator/sub_OpPsi_SG4.f90
:1488-1538) takes every component basis->grid, adds
GridOp(:, ib0, jb0, iterm) * d^{iterm} Psi(:, jb0) into OpPsi(:, ib0) over
ib0, jb0 and the KEO terms, and goes grid->basis.  Every step is linear in
the grids, so the (ib0, jb0) block of H is the single-surface operator whose
grids are GridOp(:, ib0, jb0, :) (kinetic terms on the diagonal blocks only
when, as usual, they do not couple the surfaces), and

    (H Psi)_ib0 = sum_jb0 H_{ib0 jb0} Psi_jb0.

Each block is an SG4Operator (constant-coefficient or general type 1, type
10) on the same basis, or None for a zero block; a block object given for
both (i, j) and (j, i) (a symmetric coupling) is applied once to both
components stacked in one launch.

Packed layout: a vector holds the nb0 components one after the other, nb_ba
values each (basis index fastest, then the electronic one), as RvecB of a
param_psi with nb_be = nb0 (i_bhe = i_b + nb_ba (i_e - 1) for nb_bi = 1,
Write_Psi_nDBasis, SRC/sub_WP/sub_module_psi_io.f90:1525-1530).
"""
import numpy as np
import torch

from . import PREC_FP64


class SG4Surfaces:
    """H over nb0 surfaces; ``blocks[i][j]``: SG4Operator or None."""

    def __init__(self, blocks):
        self.nb0 = len(blocks)
        if self.nb0 < 1 or any(len(r) != self.nb0 for r in blocks):
            raise ValueError("blocks must be an nb0 x nb0 list of lists")
        self.blocks = [list(r) for r in blocks]
        ops = [b for r in blocks for b in r if b is not None]
        if not ops:
            raise ValueError("at least one block must be an operator")
        self.basis = ops[0].basis
        nb = self.basis.nb_ba
        if any(o.basis.nb_ba != nb for o in ops):
            raise ValueError("every block must act on the same packed basis")
        self.max_nvec = min(o.max_nvec for o in ops)
        self._h = ops[0]._h             # stream and device for the block kernels
        self.nb_comp = nb

    @property
    def nb_ba(self):
        """rows of the multi-surface packed vector: nb0 x nb_ba"""
        return self.nb0 * self.nb_comp

    def apply(self, psi, out, prec=PREC_FP64):
        """H.psi on device tensors of shape (nvec, nb0 nb_ba); out is overwritten."""
        nb = self.nb_comp
        nvec = psi.shape[0]
        if psi.shape[1] != self.nb_ba or out.shape != psi.shape:
            raise ValueError("psi and out must be (nvec, nb0 * nb_ba)")
        comp = [psi[:, i * nb:(i + 1) * nb].contiguous() for i in range(self.nb0)]
        acc = [None] * self.nb0
        done = set()

        def add(i, y):
            acc[i] = y if acc[i] is None else acc[i].add_(y)

        for i in range(self.nb0):
            for j in range(self.nb0):
                op = self.blocks[i][j]
                if op is None or (i, j) in done:
                    continue
                if i != j and self.blocks[j][i] is op and 2 * nvec <= op.max_nvec:
                    # one launch for H_ij psi_j and H_ji psi_i (the same operator)
                    y = torch.empty(2 * nvec, nb, device=psi.device, dtype=psi.dtype)
                    op.apply(torch.cat([comp[j], comp[i]], 0), y, prec)
                    add(i, y[:nvec])
                    add(j, y[nvec:])
                    done.update({(i, j), (j, i)})
                    continue
                for a in range(0, nvec, op.max_nvec):
                    b = min(a + op.max_nvec, nvec)
                    y = torch.empty(b - a, nb, device=psi.device, dtype=psi.dtype)
                    op.apply(comp[j][a:b], y, prec)
                    if acc[i] is None:
                        acc[i] = torch.zeros(nvec, nb, device=psi.device, dtype=psi.dtype)
                    acc[i][a:b].add_(y)
                done.add((i, j))
        for i in range(self.nb0):
            if acc[i] is None:
                out[:, i * nb:(i + 1) * nb] = 0.0
            else:
                out[:, i * nb:(i + 1) * nb] = acc[i]
        return out

    def apply_host(self, psi, prec=PREC_FP64, device="cuda:0"):
        """H.psi for a host block (nb0 nb_ba,) or (nb0 nb_ba, nvec); returns numpy."""
        psi = np.asarray(psi, dtype=np.float64)
        one = psi.ndim == 1
        P = psi[:, None] if one else psi
        x = torch.tensor(np.ascontiguousarray(P.T), device=device)
        y = torch.empty_like(x)
        self.apply(x, y, prec)
        out = y.cpu().numpy().T
        return out[:, 0] if one else out

    def zeroth_order_diagonal(self, device):
        """Harmonic zeroth-order energies of every row (the Davidson's
        preconditioner diagonal), repeated for every surface."""
        from .davidson import ho_zeroth_order_diagonal
        d = ho_zeroth_order_diagonal(self.basis, device)
        return d.repeat(self.nb0)

    def set_field(self, mu, eps):
        """Only the field-free operator (mu None) is supported."""
        if mu is not None:
            raise NotImplementedError("field terms of multi-surface operators are not supported")
