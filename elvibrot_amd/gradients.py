"""Gradients of eigenvalues with respect to the potential.

For H(theta) = T + V_theta on the SG4 grids, dE/dV(p) is given by the device
kernel sg4_grad_V (W_t w(p) (B phi)(p) (B psi)(p) per grid point), and
dE/dtheta = sum_p dE/dV(p) dV(p)/dtheta.  For a non-symmetric operator (the
type-10 d1 G d1 form on a finite grid) the bra is the left eigenvector:
dE/dtheta = <l|dH/dtheta|r> / <l|r>.
"""
import numpy as np
import torch


def grad_V(op, phi, psi, device="cuda:0"):
    """phi, psi: numpy (nb_ba, n) blocks; returns numpy (sum_nq, n)."""
    dev = torch.device(device)
    n = psi.shape[1]
    out = np.zeros((op.basis.sum_nq, n))
    for a in range(0, n, op.max_nvec):
        b = min(a + op.max_nvec, n)
        P = torch.tensor(np.ascontiguousarray(psi[:, a:b]), device=dev)
        F = torch.tensor(np.ascontiguousarray(phi[:, a:b]), device=dev)
        G = torch.empty(op.basis.sum_nq, b - a, device=dev, dtype=torch.float64)
        op.grad_V(F, P, G)
        out[:, a:b] = G.cpu().numpy()
    return out


def dense_matrix(op, prec=0):
    """H as a dense host matrix built from block applications (small bases)."""
    nb = op.nb_ba
    H = np.zeros((nb, nb))
    for a in range(0, nb, op.max_nvec):
        b = min(a + op.max_nvec, nb)
        E = np.zeros((nb, b - a))
        E[np.arange(a, b), np.arange(b - a)] = 1.0
        H[:, a:b] = op.apply_host(E, prec)
    return H


def lowest_eigenpairs(H, n):
    """Lowest n eigenvalues with right and left eigenvectors (normalised so
    that <l|r> = 1); for a symmetric H, l = r."""
    if np.allclose(H, H.T, rtol=0, atol=1e-13 * np.abs(H).max()):
        E, X = np.linalg.eigh(0.5 * (H + H.T))
        return E[:n], X[:, :n], X[:, :n]
    E, R = np.linalg.eig(H)
    i = np.argsort(E.real)[:n]
    E, R = E[i].real, R[:, i].real
    El, Lv = np.linalg.eig(H.T)
    Lsel = np.zeros_like(R)
    for k, e in enumerate(E):
        j = np.argmin(np.abs(El - e))
        Lsel[:, k] = Lv[:, j].real
    Lsel /= np.sum(Lsel * R, axis=0)[None, :]
    return E, R, Lsel


def eigenvalue_gradients(op, dV_list, n, device="cuda:0", method=None, conv_resi=1e-9):
    """dE_k/dtheta_j for the n lowest levels; dV_list[j] is the grid dV/dtheta_j.

    method "davidson": eigenpairs from the device block Davidson (symmetric
    operators, any basis size: no dense matrix); "dense": the host matrix
    with left and right eigenvectors (any operator, small bases: O(nb_ba^3)).
    Default: "davidson" for the constant-coefficient type-1 operators (H
    symmetric), "dense" otherwise (type 10 on a finite grid is not)."""
    if method is None:
        method = "davidson" if getattr(op, "type_op", 1) == 1 else "dense"
    if method == "davidson":
        from .davidson import sub_propagation_Davidson
        r = sub_propagation_Davidson(op, n, max_it=500, conv_resi=conv_resi, conv_ene=1e-12,
                                     device=device)
        if not r.converged.all():
            raise RuntimeError("Davidson did not converge the requested levels")
        E = np.asarray(r.E)
        R = np.ascontiguousarray(r.X.cpu().numpy().T)
        Lv = R
    elif method == "dense":
        E, R, Lv = lowest_eigenpairs(dense_matrix(op), n)
    else:
        raise ValueError("method must be 'davidson' or 'dense'")
    g = grad_V(op, Lv, R, device)          # (sum_nq, n)
    return E, np.stack([g.T @ dV for dV in dV_list], axis=1), g
