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
the device; the small projected eigenproblem is solved on the host.  With
``prec`` = PREC_F32 the iterations run with the FP32 H.psi until every
residual is below ``switch_resi`` (above the FP32 noise floor of the
residuals); the basis is then collapsed onto the current Ritz vectors, H is
re-applied to them in FP64 and the iterations continue in FP64 to the
requested tolerances.  The FP32 phase does the bulk of the subspace
expansion; the FP64 phase (a few iterations) removes the FP32 error from the
eigenpairs.
"""
import numpy as np
import torch

from . import PREC_FP64


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
    E0 = ho_zeroth_order_diagonal(op.basis, dev)
    Vb = torch.zeros(max_WP, nb, device=dev, dtype=torch.float64)   # basis (rows = vectors)
    HV = torch.zeros_like(Vb)
    Hp = np.zeros((max_WP, max_WP))                                    # <Vb_i | H Vb_j>
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
        W = W / torch.linalg.norm(W, dim=1, keepdim=True).clamp_min(1e-300)
        for _ in range(2):   # double Gram-Schmidt against the basis, then inside the block
            if ndim:
                W = W - _gram(W, Vb[:ndim]) @ Vb[:ndim]
            lam, U = np.linalg.eigh(_gram(W, W).cpu().numpy())
            ok = lam > 1e-10
            if not ok.any():
                return None
            T = torch.tensor(U[:, ok] / np.sqrt(lam[ok]), device=dev)
            W = T.T @ W
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
            Vb[:keep] = X_all[:keep]
            HV[:keep] = HX_all[:keep]
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
        G = _gram(Vb[:ndim + m], HV[ndim:ndim + m]).cpu().numpy()
        Hp[:ndim + m, ndim:ndim + m] = G
        Hp[ndim:ndim + m, :ndim] = G[:ndim].T
        Hp[ndim:ndim + m, ndim:ndim + m] = 0.5 * (G[ndim:] + G[ndim:].T)
        ndim += m
        ev, vec = np.linalg.eigh(Hp[:ndim, :ndim])
        # Ritz vectors: the nb_WP wanted ones and those a restart keeps
        nk = min(ndim, nb_WP + block)
        vec_t = torch.tensor(np.ascontiguousarray(vec[:, :nk].T), device=dev)
        X_all = vec_t @ Vb[:ndim]
        HX_all = vec_t @ HV[:ndim]
        E = ev[:nb_WP]
        X = X_all[:nb_WP]
        R = HX_all[:nb_WP] - torch.tensor(E, device=dev)[:, None] * X
        rn = torch.linalg.norm(R, dim=1).cpu().numpy()
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
            Q = orthonormalize(X_all[:keep].clone(), 0)
            keep = Q.shape[0]
            Vb[:keep] = Q
            HV[:keep] = apply(Vb[:keep].contiguous(), PREC_FP64)
            ndim = keep
            phase_prec = PREC_FP64
            G = _gram(Vb[:ndim], HV[:ndim]).cpu().numpy()
            Hp[:ndim, :ndim] = 0.5 * (G + G.T)
            ev, vec = np.linalg.eigh(Hp[:ndim, :ndim])
            vec_t = torch.tensor(np.ascontiguousarray(vec.T), device=dev)
            X_all = vec_t @ Vb[:ndim]
            HX_all = vec_t @ HV[:ndim]
            E = ev[:nb_WP]
            X = X_all[:nb_WP]
            R = HX_all[:nb_WP] - torch.tensor(E, device=dev)[:, None] * X
            rn = torch.linalg.norm(R, dim=1).cpu().numpy()
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
