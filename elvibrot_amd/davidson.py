"""Block Davidson on the device around the SG4 H.psi library.

Mirrors the control flow of sub_propagation_Davidson
(SRC/sub_propagation/sub_module_Davidson.f90:59-763):

* each iteration applies H to the block of new vectors (sub_MakeHPsi_Davidson,
  :896-1040, block call :970), updates the projected matrix
  (sub_MakeH_Davidson, :1044-1159) and diagonalises it (:362);
* convergence of state j needs both |E_j - E_j(previous)| < conv_ene and
  ||r_j|| < conv_resi (:420-473);
* new vectors are the residuals r_j = sum_i (H psi_i - E_j psi_i) Vec(i,j)
  (MakeResidual_Davidson, :1230-1304) divided by E_j - E0(ib), with the
  reference's guard 1/(D + 1e-3) when |D| <= conv_resi (NewVec_type = 4,
  sub_NewVec_Davidson :1440-1460), then normalised and orthogonalised twice
  against the basis (:1529-1547), here as a block: two passes of block
  Gram-Schmidt (two GEMMs each) followed by an orthonormalisation inside the
  block from the eigen-decomposition of its Gram matrix, dropping directions
  whose squared norm is below 1e-10 (the reference's cut for a vector after
  its Gram-Schmidt); the span, hence the Ritz values, is the same as with the
  reference's vector-by-vector sweep;
* when the basis would exceed max_WP it is collapsed onto the current Ritz
  vectors (restart).

All block algebra (projections, residuals, Gram-Schmidt, norms) is FP64 on
the device, the products over the packed vectors in the library's block
kernels (sg4_block_gram, sg4_block_combine); the small projected
eigenproblem is solved on the host.  With
``prec`` = PREC_F32 the iterations run with the FP32 H.psi until every
residual is below ``switch_resi`` (above the FP32 noise floor of the
residuals); the basis is then collapsed onto the current Ritz vectors, H is
re-applied to them in FP64 and the iterations continue in FP64 to the
requested tolerances.  The FP32 phase does the bulk of the subspace
expansion; the FP64 phase (a few iterations) removes the FP32 error from the
eigenpairs.
"""
import ctypes as C

import numpy as np
import torch

from . import PREC_FP64, _capi
from ._capi import lib


def ho_zeroth_order_diagonal(basis, device):
    """E0(ib) = sum_k (i_k + 1/2) scaleQ-independent HO energies of the packed
    basis functions (the role of get_Ene0_AT_ib_FROM_Basis,
    SRC/sub_Basis/sub_module_basis_set_alloc.f90:3122, for HO primitive bases)."""
    idx = basis.packed()
    e0 = idx.sum(axis=1) + 0.5 * basis.D
    return torch.tensor(e0, device=device, dtype=torch.float64)


def _gram(A, B, chunks=256):
    """A @ B.T for row blocks A (m, n), B (p, n) with a long inner dimension
    (n = nb_ba): one batched product over `chunks` pieces of n (strided views
    of A and B, no copies) summed over the pieces, plus the remainder.  For
    small m, p and n ~ 1e6 the library GEMM picks kernels without a K split:
    A @ B.T takes 2.4 ms for 84 x 84 and 26 ms for 64 x 1 at n = 1.26e6, this
    form 0.3-0.6 ms for every shape up to 84 x 84."""
    m, n = A.shape
    p = B.shape[0]
    k = n // chunks
    n0 = k * chunks
    out = None
    if k > 0:
        Ac = A[:, :n0].reshape(m, chunks, k).transpose(0, 1)      # (chunks, m, k)
        Bc = B[:, :n0].reshape(p, chunks, k).permute(1, 2, 0)     # (chunks, k, p)
        out = torch.bmm(Ac, Bc).sum(0)
    if n0 < n:
        r = A[:, n0:] @ B[:, n0:].T
        out = r if out is None else out + r
    return out


class _Dense:
    """Products over packed vectors (rows of length nb_ba) on the device.
    Grams with at most 400 outputs or at most 16 columns (up to 124 rows)
    and rotations to at most 20 rows go through the library's block kernels
    (sg4_block_gram, sg4_block_combine, on the operator's plan and stream);
    the others through the split-K batched product and the library GEMM,
    which are faster there (10D L=8: gram 20 x 20 0.19 against 0.32 ms,
    100 x 16 0.47 against 0.57 ms, 16 x 100 0.56 against 0.31 ms, 36 x 36
    0.37 against 0.30 ms; rotation 20 x 84 0.34 against 0.59 ms, 36 x 84
    0.69 against 0.64 ms).
    CPU tensors (host tests of the driver with a dense matrix) use plain
    torch products."""

    def __init__(self, op):
        self.h = getattr(op, "_h", None)
        self.work = None

    def _rows(self, A):
        return A if A.stride(1) == 1 else A.contiguous()

    def gram(self, A, B):
        """A @ B.T"""
        if A.device.type != "cuda" or self.h is None:
            return _gram(A, B)
        A, B = self._rows(A), self._rows(B)
        m, n = A.shape
        q = B.shape[0]
        need = C.c_int64()
        if (m * q > 400 and q > 16) or lib().sg4_block_gram_work(m, q, C.byref(need)) != 0:
            return _gram(A, B)
        if self.work is None or self.work.numel() < need.value:
            self.work = torch.empty(need.value, device=A.device, dtype=torch.float64)
        out = torch.empty(m, q, device=A.device, dtype=torch.float64)
        _capi.check(lib().sg4_block_gram(self.h, m, q, n, C.c_void_p(A.data_ptr()), A.stride(0),
                                         C.c_void_p(B.data_ptr()), B.stride(0),
                                         C.c_void_p(out.data_ptr()), C.c_void_p(self.work.data_ptr()),
                                         self.work.numel()))
        return out

    def comb(self, U, X, Y=None, beta=0.0):
        """beta Y + U @ X (into Y when given; U: k x d, X: d rows)"""
        U = torch.as_tensor(U, dtype=torch.float64, device=X.device)
        k, d = U.shape
        if X.device.type != "cuda" or self.h is None or k > 20 or k * d > 8192:
            R = U @ X
            if Y is None:
                return R
            Y.mul_(beta).add_(R) if beta != 0.0 else Y.copy_(R)
            return Y
        X = self._rows(X)
        U = U.contiguous()
        if Y is None:
            Y = torch.empty(k, X.shape[1], device=X.device, dtype=torch.float64)
            beta = 0.0
        _capi.check(lib().sg4_block_combine(self.h, k, d, X.shape[1], C.c_void_p(U.data_ptr()),
                                            C.c_void_p(X.data_ptr()), X.stride(0), float(beta),
                                            C.c_void_p(Y.data_ptr()), Y.stride(0)))
        return Y


class DavidsonResult:
    def __init__(self, E, X, it, n_hpsi, converged, E_fp64=None, history=None):
        self.E, self.X, self.it, self.n_hpsi = E, X, it, n_hpsi
        self.converged, self.E_fp64, self.history = converged, E_fp64, history or []


def sub_propagation_Davidson(op, nb_WP, max_it=200, max_WP=None, conv_resi=5e-4,
                             conv_ene=1e-4, block=None, prec=PREC_FP64, psi0=None,
                             final_rr_fp64=True, device="cuda:0", verbose=False,
                             switch_resi=None, refine=True):
    """Lowest ``nb_WP`` eigenpairs of the SG4 operator ``op``.

    Returns a DavidsonResult with E (host numpy), X (device, nb_WP x nb_ba),
    iterations, number of H.psi vector applications and, when ``prec`` is not
    FP64, the FP64 Rayleigh-Ritz eigenvalues E_fp64 (E is then E_fp64).
    ``refine=False`` keeps the reduced precision to the end (no FP64 phase):
    the eigenvalues are then those of the reduced-precision operator, the run
    stopping at convergence or when the residuals stall at its noise floor."""
    nb = op.nb_ba
    block = block or min(op.max_nvec, nb_WP)
    max_WP = max_WP or max(4 * nb_WP, nb_WP + 4 * block)
    max_WP = min(max_WP, nb)
    dev = torch.device(device)
    E0 = (op.zeroth_order_diagonal(dev) if hasattr(op, "zeroth_order_diagonal")
          else ho_zeroth_order_diagonal(op.basis, dev))
    Vb = torch.zeros(max_WP, nb, device=dev, dtype=torch.float64)   # basis (rows = vectors)
    HV = torch.zeros_like(Vb)
    Hp = np.zeros((max_WP, max_WP))                                    # <Vb_i | H Vb_j>
    dn = _Dense(op)
    if psi0 is None:
        order = torch.argsort(E0)[:nb_WP]
        psi0 = torch.zeros(nb_WP, nb, device=dev, dtype=torch.float64)
        psi0[torch.arange(nb_WP, device=dev), order] = 1.0
    new = psi0
    ndim = 0
    n_hpsi = 0
    Ene0 = None
    E = X = None
    conv = np.zeros(nb_WP, dtype=bool)
    hist = []

    def apply(blk, p):
        nonlocal n_hpsi
        out = torch.empty_like(blk)
        for a in range(0, blk.shape[0], op.max_nvec):
            b = min(a + op.max_nvec, blk.shape[0])
            op.apply(blk[a:b].contiguous(), out[a:b], p)
        n_hpsi += blk.shape[0]
        return out

    def orthonormalize(W, ndim):
        W = W.contiguous()
        nrm = torch.sqrt(torch.diagonal(dn.gram(W, W))).clamp_min(1e-300)
        W = W / nrm[:, None]
        for _ in range(2):   # double Gram-Schmidt against the basis, then inside the block
            if ndim:
                dn.comb(-dn.gram(W, Vb[:ndim]), Vb[:ndim], W, 1.0)
            lam, U = np.linalg.eigh(dn.gram(W, W).cpu().numpy())
            ok = lam > 1e-10
            if not ok.any():
                return None
            W = dn.comb(np.ascontiguousarray((U[:, ok] / np.sqrt(lam[ok])).T), W)
        return W

    phase_prec = prec
    # the FP32 residual floor is ~ eps32 x (Smolyak roundoff amplification) x ||H||
    switch_resi = max(conv_resi, 1e-3) if switch_resi is None else switch_resi
    best_low = np.inf
    stall = 0
    n_low = 0
    it = 0
    for it in range(max_it + 1):
        W = orthonormalize(new, ndim)
        if W is None:
            break
        m = W.shape[0]
        if ndim + m > max_WP:          # restart on the current Ritz vectors
            keep = min(max(nb_WP + block, nb_WP), max_WP - m)
            vk = np.ascontiguousarray(vec[:, :keep].T)
            Xk, HXk = dn.comb(vk, Vb[:ndim]), dn.comb(vk, HV[:ndim])
            Vb[:keep] = Xk
            HV[:keep] = HXk
            Hp[:keep, :keep] = vec[:, :keep].T @ Hp[:ndim, :ndim] @ vec[:, :keep]
            ndim = keep
            W = orthonormalize(new, ndim)
            if W is None:
                break
            m = W.shape[0]
        Vb[ndim:ndim + m] = W
        HV[ndim:ndim + m] = apply(W, phase_prec)
        # projected matrix: only the columns of the new vectors are computed,
        # Hp[i, j] = <Vb_i | H Vb_j> for i < j mirrored (H symmetric), the new
        # diagonal block symmetrised
        G = dn.gram(Vb[:ndim + m], HV[ndim:ndim + m]).cpu().numpy()
        Hp[:ndim + m, ndim:ndim + m] = G
        Hp[ndim:ndim + m, :ndim] = G[:ndim].T
        Hp[ndim:ndim + m, ndim:ndim + m] = 0.5 * (G[ndim:] + G[ndim:].T)
        ndim += m
        ev, vec = np.linalg.eigh(Hp[:ndim, :ndim])
        # the nb_WP wanted Ritz vectors (those a restart keeps beyond them are
        # formed at the restart, from vec)
        vec_t = np.ascontiguousarray(vec[:, :nb_WP].T)
        X = dn.comb(vec_t, Vb[:ndim])
        HX = dn.comb(vec_t, HV[:ndim])
        E = ev[:nb_WP]
        R = HX - torch.tensor(E, device=dev)[:, None] * X
        rn = np.sqrt(np.maximum(np.diagonal(dn.gram(R, R).cpu().numpy()), 0.0))
        dE = np.abs(E - Ene0) if Ene0 is not None else np.full(nb_WP, np.inf)
        conv = (rn < conv_resi) & (dE < conv_ene)
        hist.append((it, ndim, float(rn.max()), float(dE.max())))
        if verbose:
            print(f"davidson it {it} ndim {ndim} max|r| {rn.max():.2e} max|dE| {dE.max():.2e} conv {conv.sum()}")
        Ene0 = E.copy()
        if phase_prec != PREC_FP64:
            stall = stall + 1 if rn.max() > 0.7 * best_low else 0
            best_low = min(best_low, rn.max())
        if phase_prec != PREC_FP64 and not refine:
            # no FP64 phase: run to convergence or until the residuals have
            # not improved for 20 iterations (the reduced precision's floor)
            if (conv.all() and it > 0) or stall >= 20:
                break
        elif phase_prec != PREC_FP64 and (rn.max() < switch_resi or stall >= 6):
            # FP64 phase: restart on the Ritz vectors with an FP64 H.psi
            n_low = n_hpsi
            keep = min(ndim, nb_WP + block)
            Q = orthonormalize(dn.comb(np.ascontiguousarray(vec[:, :keep].T), Vb[:ndim]), 0)
            keep = Q.shape[0]
            Vb[:keep] = Q
            HV[:keep] = apply(Vb[:keep].contiguous(), PREC_FP64)
            ndim = keep
            phase_prec = PREC_FP64
            G = dn.gram(Vb[:ndim], HV[:ndim]).cpu().numpy()
            Hp[:ndim, :ndim] = 0.5 * (G + G.T)
            ev, vec = np.linalg.eigh(Hp[:ndim, :ndim])
            vec_t = np.ascontiguousarray(vec[:, :nb_WP].T)
            X = dn.comb(vec_t, Vb[:ndim])
            HX = dn.comb(vec_t, HV[:ndim])
            E = ev[:nb_WP]
            R = HX - torch.tensor(E, device=dev)[:, None] * X
            rn = np.sqrt(np.maximum(np.diagonal(dn.gram(R, R).cpu().numpy()), 0.0))
            conv = np.zeros(nb_WP, dtype=bool)
            Ene0 = E.copy()
        elif conv.all() and phase_prec == PREC_FP64:
            break
        todo = np.nonzero(~conv)[0][:block]
        Rt = R[todo]
        Et = torch.tensor(E[todo], device=dev)[:, None]
        Dg = Et - E0[None, :]
        Dg = torch.where(Dg.abs() > conv_resi, Dg, Dg + 1e-3)
        new = Rt / Dg
    E_fp64 = E if (prec != PREC_FP64 and refine) else None
    res = DavidsonResult(E, X, it, n_hpsi, conv, E_fp64, hist)
    res.n_hpsi_low = n_low      # H.psi vector applications done in the reduced precision
    return res
