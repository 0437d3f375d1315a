"""The kinetic part of the Smolyak SG4 type-1 operator as one sparse matrix.

For 1-D bases whose quadrature integrates products of two basis functions
exactly (B^T W B = I: HO and Legendre bases with nq >= nb), the kinetic part
of the combination sum  sum_t W_t P_t^T (sum_k c_k B^T W D2_k B) P_t  reduces,
because the combination weights telescope, to

    K[r][r'] = c_k T_k(L - Lambda_k(r))[i_k][i'_k],   T_k(m) = BtW D2 B at level m,

for packed rows r, r' differing in mode k only (Lambda_k(r) = sum over the
other modes of the smallest level holding their index).  The device split
path (sg4_plan_split_info) applies this K.  These tests build K in numpy
from the oracle's tables and check it against the oracle's own term-by-term
Smolyak sum (SRC/sub_Operator/sub_OpPsi_SG4.f90:681-977 restated in
oracle/sg4_oracle.c): oracle H(ck) - H(ck = 0) is the brute-force kinetic
part, so a wrong level (L - Lambda off by one), a dropped mode or a
transposed T fails here.
"""
import numpy as np
import pytest

from oracle import oracle as O


def split_kinetic(o, ck):
    """Dense K of the closed form (small cases only)."""
    D, nb_ba = o.D, o.nb_ba
    L = int(np.max(o.lev.sum(1)))
    nbtab = np.array([[o.basis1d(k, l)["nb"] for l in range(o.L + 1)] for k in range(D)])
    lam = [np.array([int(np.argmax(nbtab[k] > i)) for i in range(nbtab[k, -1])]) for k in range(D)]
    rows = {tuple(r): n for n, r in enumerate(o.packed)}
    T = {}
    K = np.zeros((nb_ba, nb_ba))
    for n, r in enumerate(o.packed):
        lr = np.array([lam[k][r[k]] for k in range(D)])
        for k in range(D):
            m = L - (lr.sum() - lr[k])
            if (k, m) not in T:
                b = o.basis1d(k, m)
                d0, w, g2 = b["d0"], b["w"], b["gg2"]
                T[k, m] = d0.T @ (w[:, None] * (g2 @ d0))
            Tk = T[k, m]
            for i2 in range(Tk.shape[0]):
                r2 = list(r)
                r2[k] = i2
                n2 = rows.get(tuple(r2))
                if n2 is not None:
                    K[n, n2] += ck[k] * Tk[r[k], i2]
    return K


@pytest.mark.parametrize("D,L,LB,nq_A,nb_A", [(2, 6, None, 1, None), (3, 4, None, 1, None),
                                              (4, 3, None, 1, None), (3, 5, 3, 1, None),
                                              (3, 4, None, 3, 2)])
def test_kinetic_closed_form_matches_smolyak_sum(D, L, LB, nq_A, nb_A):
    o = O.SG4(D, L, LB=LB, nq_A=nq_A, nb_A=nb_A)
    rng = np.random.default_rng(7)
    ck = -0.5 * (1.0 + 0.3 * rng.standard_normal(D))
    V = rng.standard_normal(o.sum_nq)
    psi = rng.standard_normal((o.nb_ba, 3))
    kin = o.hpsi(psi, V, ck) - o.hpsi(psi, V, np.zeros(D))
    K = split_kinetic(o, ck)
    got = K @ psi
    err = np.max(np.abs(got - kin)) / np.max(np.abs(kin))
    print(f"D={D} L={L} LB={LB}: nb_ba {o.nb_ba}, nnz {np.count_nonzero(K)}, max rel {err:.2e}")
    assert err < 1e-12


def test_kinetic_matrix_is_symmetric_and_level_dependent():
    """K is symmetric (T_k(m) is, for an orthonormal basis and exact
    quadrature of the products), and rows with higher excitations in the
    other modes use lower-level matrices: dropping the level dependence
    (T_k(L) everywhere) changes the operator."""
    o = O.SG4(3, 4)
    ck = -0.5 * np.ones(3)
    K = split_kinetic(o, ck)
    assert np.max(np.abs(K - K.T)) < 1e-12 * np.max(np.abs(K))
    V = np.zeros(o.sum_nq)
    psi = np.random.default_rng(3).standard_normal((o.nb_ba, 2))
    kin = o.hpsi(psi, V, ck)
    wrong = np.zeros_like(K)
    rows = {tuple(r): n for n, r in enumerate(o.packed)}
    for k in range(3):
        b = o.basis1d(k, 4)
        Tk = b["d0"].T @ (b["w"][:, None] * (b["gg2"] @ b["d0"]))
        for n, r in enumerate(o.packed):
            for i2 in range(Tk.shape[0]):
                r2 = list(r)
                r2[k] = i2
                n2 = rows.get(tuple(r2))
                if n2 is not None:
                    wrong[n, n2] += ck[k] * Tk[r[k], i2]
    assert np.max(np.abs(wrong @ psi - kin)) > 1e-6 * np.max(np.abs(kin))
