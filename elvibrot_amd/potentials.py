"""Operator grids of the SG4 configurations (what get_OpGrid_type1/10_OF_ONEDP_
FOR_SG4 evaluates on the first call, SRC/sub_Operator/sub_OpPsi_SG4.f90
:2536-3082).

* Henon-Heiles: evaluated on the device (SG4Operator(..., ("henon_heiles", lam))).
* Triatomic in Jacobi coordinates (R, r, theta), J = 0, type-10 operator
  (evaluated on the device by default, on the host by jacobi_triatomic_grids):
  G = diag(1/mu_R, 1/mu_r, 1/(mu_R R^2) + 1/(mu_r r^2)), Jac = sin(theta),
  sqRhoOVERJac = 1 (the basis is orthonormal for dR dr sin(theta) dtheta), and
  V = Morse(R) + Morse(r) + c cos^2(theta).
"""
import numpy as np

from . import SG4Basis, SG4Operator, BASIS_HO, BASIS_LEGENDRE

AMU = 1822.888486209  # electron masses per unified atomic mass unit


def morse(x, D, a, xe):
    return D * (1.0 - np.exp(-a * (x - xe))) ** 2


def jacobi_triatomic_basis(L, mu_R=AMU, mu_r=AMU, D=0.2, a=1.0, Re=3.0, re=2.0, nA=2, nB=2):
    """HO bases for R and r centred at the Morse minima with the Morse
    harmonic frequency (scaleQ = sqrt(mu omega), omega = a sqrt(2D/mu)),
    Legendre basis for theta; nq(l) = nb(l) = nA + nB l for all three.

    nA = 2 by default: the d1 G d1 (type-10) form needs two points on the
    lowest level, since the first-derivative matrix of a one-point rule is
    zero and the Smolyak combination would then drop the kinetic energy of
    every term with a level-0 mode."""
    sR = np.sqrt(mu_R * a * np.sqrt(2 * D / mu_R))
    sr = np.sqrt(mu_r * a * np.sqrt(2 * D / mu_r))
    n = dict(nq_A=nA, nq_B=nB, nb_A=nA, nb_B=nB)
    modes = [dict(kind=BASIS_HO, Q0=Re, scaleQ=sR, **n), dict(kind=BASIS_HO, Q0=re, scaleQ=sr, **n),
             dict(kind=BASIS_LEGENDRE, **n)]
    return SG4Basis(3, L, L, modes)


def jacobi_triatomic_grids(basis, mu_R=AMU, mu_r=AMU, D=0.2, a=1.0, Re=3.0, re=2.0,
                           coupling=0.01):
    """V, G (D x D per point, G[i + 3 j + 9 p]), Jac, sqRJ on every term grid."""
    q = np.concatenate([basis.term_grid(t) for t in range(basis.nterms)])
    R, r, th = q[:, 0], q[:, 1], q[:, 2]
    V = morse(R, D, a, Re) + morse(r, D, a, re) + coupling * np.cos(th) ** 2
    G = np.zeros((q.shape[0], 3, 3))
    G[:, 0, 0] = 1.0 / mu_R
    G[:, 1, 1] = 1.0 / mu_r
    G[:, 2, 2] = 1.0 / (mu_R * R ** 2) + 1.0 / (mu_r * r ** 2)
    Jac = np.sin(th)
    sqRJ = np.ones_like(th)
    return V, G.reshape(-1), Jac, sqRJ


def jacobi_triatomic_operator(L=10, max_nvec=8, device_grids=True, mu_R=AMU, mu_r=AMU, D=0.2,
                              a=1.0, Re=3.0, re=2.0, coupling=0.01):
    """The type-10 operator; its grids are evaluated on the device
    (sg4_plan_set_jacobi_triatomic) or, with device_grids=False, on the host
    (jacobi_triatomic_grids) and uploaded."""
    b = jacobi_triatomic_basis(L, mu_R=mu_R, mu_r=mu_r, D=D, a=a, Re=Re, re=re)
    if device_grids:
        return SG4Operator(b, max_nvec=max_nvec,
                           potential=("jacobi_triatomic", (mu_R, mu_r, D, a, Re, re, coupling)))
    V, G, Jac, sqRJ = jacobi_triatomic_grids(b, mu_R, mu_r, D, a, Re, re, coupling)
    return SG4Operator(b, max_nvec=max_nvec, potential=V, type_op=10, G=G, Jac=Jac, sqRJ=sqRJ)
