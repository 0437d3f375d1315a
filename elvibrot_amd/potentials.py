"""Operator grids of the SG4 configurations (what get_OpGrid_type1/10_OF_ONEDP_
FOR_SG4 evaluates on the first call, SRC/sub_Operator/sub_OpPsi_SG4.f90
:2536-3082).

* Henon-Heiles: evaluated on the device (SG4Operator(..., ("henon_heiles", lam))).
* Two-surface linear vibronic-coupling model of pyrazine in 12 or 24
  dimensionless normal modes (the reference's propagation tests,
  Working_tests/MPI_tests/{12D,24D}_propagation_S3: sub_system_pyrazine.f
  calcN_op and calc_f2_f1Q.f90): pyrazine_surfaces.
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


# ---- pyrazine (two surfaces, 12 or 24 modes) ----------------------------------
# Model parameters in eV (frequencies w, intra-state gradients k1, k2 of the
# two diabatic states, gap 2 delta, inter-state coupling lambda along mode 1):
# the reference tests' tables (sub_system_pyrazine.f of 12D_propagation_S3 and
# 24D_propagation_S3), kept as they are since the stored autocorrelations are
# those of these models.  k2 = (0, 0.1194, 0.2012, 0.0484, -k1[4:]).
_PYR_W = {
    12: [0.09357, 0.0740, 0.1273, 0.1568, 0.1347, 0.3431, 0.1157, 0.3242, 0.3621, 0.2673, 0.3052,
         0.0968],
    24: [0.09357, 0.0740, 0.1273, 0.1568, 0.0400, 0.0589, 0.0778, 0.0968, 0.1157, 0.1347, 0.1536,
         0.1726, 0.1915, 0.2105, 0.2294, 0.2484, 0.2673, 0.2863, 0.3052, 0.3242, 0.3431, 0.3621,
         0.3810, 0.4000]}
_PYR_K1 = {
    12: [0.0, -0.0964, 0.0470, 0.1594, 0.0308, 0.0782, 0.0261, 0.0717, 0.0560, 0.0625, 0.0780,
         0.0188],
    24: [0.0, -0.0964, 0.0470, 0.1594, 0.0069, 0.0112, 0.0102, 0.0188, 0.0261, 0.0308, 0.0210,
         0.0265, 0.0196, 0.0281, 0.0284, 0.0361, 0.0560, 0.0433, 0.0625, 0.0717, 0.0782, 0.0780,
         0.0269, 0.0306]}
PYRAZINE_DELTA, PYRAZINE_LAMBDA = 0.46165, 0.1825
# nq(l) = nb(l) = 1 + B_k l of the tests (L_TO_nq_B)
_PYR_B = {12: (3, 3, 2, 2) + (2,) * 8, 24: (4, 4, 3, 3) + (2,) * 20}


def pyrazine_model(D):
    """(w, k1, k2) in eV of the D = 12 or 24 mode model."""
    w, k1 = np.array(_PYR_W[D]), np.array(_PYR_K1[D])
    k2 = np.concatenate([[0.0, 0.1194, 0.2012, 0.0484], -k1[4:]])
    return w, k1, k2


def codata2006_au():
    """(eV per hartree, seconds per atomic unit of time) from the CODATA 2006
    constants, the reference's default set (Ext_Lib/ConstPhys_loc/SRC/
    sub_module_constant.f90, version CODATA2006): Eh = e^2 / (4 pi eps0 a0),
    a0 = 4 pi eps0 hbar^2 / (me e^2), Ta = hbar / Eh."""
    c, mu0 = 299792458.0, 4e-7 * np.pi
    h, e, me = 6.62606896e-34, 1.602176487e-19, 9.10938215e-31
    eps0, hb = 1.0 / (mu0 * c * c), h / (2 * np.pi)
    a0 = (hb / e) ** 2 / me * 4 * np.pi * eps0
    Eh = (e * e / a0) / (4 * np.pi * eps0)
    return Eh / e, hb / Eh


def pyrazine_basis(D=12, L=1):
    """HO bases in the dimensionless normal coordinates (Q0 = 0, scaleQ = 1),
    nq(l) = nb(l) = 1 + B_k l per mode (the tests' L_TO_nq_B)."""
    return SG4Basis(D, L, L, [dict(kind=BASIS_HO, nq_A=1, nq_B=b) for b in _PYR_B[D]])


def pyrazine_surfaces(basis, max_nvec=16):
    """SG4Surfaces of the model: T = -1/2 sum w_k d^2/dQ_k^2 on both states,
    V11 = sum w Q^2/2 - delta + k1.Q, V22 = sum w Q^2/2 + delta + k2.Q,
    V12 = lambda Q_1, all in hartree (calcN_op of sub_system_pyrazine.f,
    Tdef2 of calc_f2_f1Q.f90)."""
    from .surfaces import SG4Surfaces
    w, k1, k2 = pyrazine_model(basis.D)
    ev = 1.0 / codata2006_au()[0]
    Q = np.concatenate([basis.term_grid(t) for t in range(basis.nterms)])
    harm = 0.5 * (Q * Q) @ w
    ck = -0.5 * w * ev
    H1 = SG4Operator(basis, max_nvec, potential=(harm - PYRAZINE_DELTA + Q @ k1) * ev, ck=ck)
    H2 = SG4Operator(basis, max_nvec, potential=(harm + PYRAZINE_DELTA + Q @ k2) * ev, ck=ck)
    K = SG4Operator(basis, max_nvec, potential=PYRAZINE_LAMBDA * Q[:, 0] * ev, ck=np.zeros(basis.D))
    return SG4Surfaces([[H1, K], [K, H2]])
