"""The oracle pinned against the reference's own fixtures and against the
mathematics (CPU only).

Fixtures (tests/golden/, copied by fetch_reference_fixtures.sh):
  hermquadra/herm{n}.txt  -- Internal_data/HermQuadra (x, w, w e^{x^2}, 50 digits)
  hh6d_L3_benchmark.txt   -- Working_tests/MPI_tests/6D_Davidson_S0/benchmark
  hh21d_L2_benchmark.txt  -- Working_tests/MPI_tests/21D_Davidson_S0/benchmark
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O


def read_herm(golden, n):
    rows = [l.split() for l in open(os.path.join(golden, "hermquadra", f"herm{n}.txt")) if l.strip()]
    x = np.array([float(r[1]) for r in rows])
    w4 = np.array([float(r[3]) for r in rows])
    return x, w4


@pytest.mark.parametrize("n", list(range(1, 41)))
def test_gauss_hermite_vs_reference_tables(golden, n):
    # sub_quadra_herm.f90:160-178 overwrites the nodes/weights with these files
    x, w = O.gauss_hermite(n)
    xr, wr = read_herm(golden, n)
    assert np.max(np.abs(x - xr)) <= 4e-15 * max(1.0, np.max(np.abs(xr)))
    assert np.max(np.abs(w - wr) / wr) <= 1e-14


def test_gauss_legendre_exactness():
    for n in (1, 2, 5, 11, 21):
        x, w = O.gauss_legendre(n)
        for deg in range(2 * n):
            exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
            assert abs(np.sum(w * x ** deg) - exact) < 1e-14


@pytest.mark.parametrize("nq", [1, 3, 9, 17, 25])
def test_ho_basis_orthonormal_and_ode(nq):
    x, w, d0, d1, d2 = O.ho_basis(nq, nq)
    S = d0.T @ (w[:, None] * d0)
    assert np.max(np.abs(S - np.eye(nq))) < 1e-13
    # Hermite ODE: phi_n'' = (x^2 - 2n - 1) phi_n
    n = np.arange(nq)
    assert np.max(np.abs(d2 - (x[:, None] ** 2 - 2 * n - 1) * d0)) < 1e-12
    # ladder identity phi_n' = sqrt(n/2) phi_{n-1} - sqrt((n+1)/2) phi_{n+1}
    for m in range(nq - 1):
        ref = -np.sqrt((m + 1) / 2) * d0[:, m + 1]
        if m > 0:
            ref = ref + np.sqrt(m / 2) * d0[:, m - 1]
        assert np.max(np.abs(d1[:, m] - ref)) < 1e-12


def test_ho_scaling():
    # Q = Q0 + x/scaleQ; basis stays orthonormal on the scaled weights,
    # derivatives scale as sqrt(s) s^m (sub_scale_basis, sub_module_basis.f90:1499)
    x0, w0, a0, a1, a2 = O.ho_basis(7, 7)
    x, w, b0, b1, b2 = O.ho_basis(7, 7, Q0=0.5, scaleQ=2.0)
    assert np.allclose(x, 0.5 + x0 / 2.0)
    assert np.allclose(b1, a1 * 2 ** 1.5) and np.allclose(b2, a2 * 2 ** 2.5)
    assert np.max(np.abs(b0.T @ (w[:, None] * b0) - np.eye(7))) < 1e-13


@pytest.mark.parametrize("nq", [3, 9, 17])
def test_grid_derivative_matrices_exact_on_basis(nq):
    x, w, d0, d1, d2 = O.ho_basis(nq, nq)
    g0, g1, g2 = O.grid_deriv(d0, d1, d2)
    # dnRGG%dk applied to the grid values of a basis function gives its
    # derivative values (the Check block of sub_dnGB_TO_dnGG)
    assert np.max(np.abs(g0 - np.eye(nq))) < 1e-12
    assert np.max(np.abs(g1 @ d0 - d1)) < 1e-11
    assert np.max(np.abs(g2 @ d0 - d2)) < 1e-11


def test_grid_derivative_rectangular():
    x, w, d0, d1, d2 = O.ho_basis(9, 5)
    g0, g1, g2 = O.grid_deriv(d0, d1, d2)
    assert np.max(np.abs(g1 @ d0 - d1)) < 1e-11
    assert np.max(np.abs(g2 @ d0 - d2)) < 1e-11


@pytest.mark.parametrize("D,L,nterms,nb_ba,sum_nq", [
    (6, 3, 84, 377, 1216), (6, 6, 923, 8989, 93535), (21, 2, 253, 925, 2059),
    (3, 10, 166, None, None)])
def test_smolyak_counts_and_weights(D, L, nterms, nb_ba, sum_nq):
    s = O.SG4(D, L)
    Lmin = max(0, L - D + 1)
    # number of terms: sum_{m=Lmin}^{L} C(m+D-1, D-1)
    n = sum(math.comb(m + D - 1, D - 1) for m in range(Lmin, L + 1))
    assert s.nterms == n
    if nterms is not None and nb_ba is not None:
        assert s.nterms == nterms and s.nb_ba == nb_ba and s.sum_nq == sum_nq
    # the combination technique reproduces constants: sum of weights is 1
    assert abs(s.weight.sum() - 1.0) < 1e-12
    # packed basis count from the generating function of lambda(i)=ceil((i-1)/2)
    poly = np.zeros(L + 1, dtype=object); poly[0] = 1
    c = [1] + [2] * L
    for _ in range(D):
        new = np.zeros(L + 1, dtype=object)
        for a in range(L + 1):
            for b in range(L + 1 - a):
                new[a + b] += poly[a] * c[b]
        poly = new
    assert s.nb_ba == int(poly.sum())
    # every packed function is reached by at least one term, map rows valid
    assert s.map.min() >= 0 and s.map.max() == s.nb_ba - 1
    assert len(np.unique(s.map)) == s.nb_ba


def test_term_order_first_index_fastest():
    s = O.SG4(3, 2)
    # ADD_ONE_TO_nDindex_type5m order for D=3, L=2, Lmin=0
    expect = [(0,0,0),(1,0,0),(2,0,0),(0,1,0),(1,1,0),(0,2,0),(0,0,1),(1,0,1),(0,1,1),(0,0,2)]
    assert [tuple(r) for r in s.lev] == expect


def test_packed_order_last_index_fastest():
    s = O.SG4(2, 1)
    # lambda(i)=ceil((i-1)/2) (1-based): index 0 -> 0, 1,2 -> 1 ; sum <= 1
    expect = [(0, 0), (0, 1), (0, 2), (1, 0), (2, 0)]
    assert [tuple(r) for r in s.packed] == expect


def test_smolyak_exact_for_separable_polynomial_potential():
    # For V = sum_k a_k q_k^2 (separable quadratic) the SG4 matrix of a
    # tensor basis is exact where every 1-D factor is exactly integrated; the
    # diagonal of the operator on the lowest function is sum_k (a_k/2 + 1/4).
    s = O.SG4(4, 3)
    a = np.array([0.5, 0.7, 1.1, 0.3])
    V = np.concatenate([(s.term_grid(t) ** 2 @ a) for t in range(s.nterms)])
    e0 = np.zeros(s.nb_ba); e0[0] = 1.0
    h = s.hpsi(e0, V, -0.5 * np.ones(4))
    assert abs(h[0] - np.sum(a / 2 + 0.25)) < 1e-12


@pytest.mark.parametrize("D,L", [(1, 0), (1, 3), (1, 6), (1, 15), (3, 0), (5, 0)])
def test_smallest_sets_closed_forms(D, L):
    """One mode: the only term is l = (L) with weight 1 and the Henon-Heiles
    potential is q^2/2, so H is the harmonic oscillator in its own basis,
    diag(n + 1/2), exact because the 2L+1 Gauss-Hermite points integrate
    every product the matrix needs.  L = 0: one function (0,...,0) on the
    one-point grid q = 0, where V = 0 and -phi0''/2 = phi0/2 per mode: D/2."""
    s = O.SG4(D, L)
    assert s.nterms == 1
    H = s.hpsi(np.eye(s.nb_ba), s.grid_hh(0.111803), -0.5 * np.ones(D))
    if L == 0:
        assert H.shape == (1, 1) and abs(H[0, 0] - D / 2) < 1e-14
    else:
        assert s.nb_ba == 2 * L + 1
        assert np.max(np.abs(H - np.diag(np.arange(s.nb_ba) + 0.5))) < 1e-12


def _dense_H(s, V, ck):
    H = s.hpsi(np.eye(s.nb_ba), V, ck)
    return 0.5 * (H + H.T), np.max(np.abs(H - H.T))


@pytest.mark.parametrize("D,L,fname", [(6, 3, "hh6d_L3_benchmark.txt"),
                                       (21, 2, "hh21d_L2_benchmark.txt")])
def test_henon_heiles_eigenvalues_vs_reference_benchmark(golden, D, L, fname):
    """The reference's Davidson levels (absolute tolerance 1e-8 in its own
    checker, Working_tests/MPI_tests/MPI_test.sh:39-41) are upper bounds of the
    exact eigenvalues of the same discrete operator, converged to
    conv_resi=5e-4 (shell_run:160-165): the exact eigenvalues must lie below
    the benchmark by at most 1e-7 and never above it by more than 1e-10."""
    ref = np.loadtxt(os.path.join(golden, fname))[:, 0]
    s = O.SG4(D, L)
    V = s.grid_hh(0.111803)
    H, asym = _dense_H(s, V, -0.5 * np.ones(D))
    assert asym < 1e-13
    ev = np.linalg.eigvalsh(H)[: len(ref)]
    d = ev - ref
    assert d.max() < 1e-10, d
    assert d.min() > -1.5e-7, d
    assert abs(d[0]) < 1e-7


def test_wrong_lambda_sign_is_caught(golden):
    # a plausible mistake (sign of the cubic coupling) moves the levels far
    # beyond the pin above; the symmetric-in-q term alone cannot hide it
    ref = np.loadtxt(os.path.join(golden, "hh6d_L3_benchmark.txt"))[:, 0]
    s = O.SG4(6, 3)
    q = [s.term_grid(t) for t in range(s.nterms)]
    lam = 0.111803
    Vbad = np.concatenate([0.5 * (g ** 2).sum(1) + lam * (g[:, :-1] ** 2 * g[:, 1:] + g[:, 1:] ** 3 / 3).sum(1) for g in q])
    H, _ = _dense_H(s, Vbad, -0.5 * np.ones(6))
    ev = np.linalg.eigvalsh(H)[:28]
    assert np.max(np.abs(ev - ref)) > 1e-4


def test_type10_constant_G_matches_type1_spectrum():
    # d1 G d1 with G = identity and Jac = sqRJ = 1 is the same continuous
    # operator as type 1; on the grid they differ by quadrature error only,
    # so the low eigenvalues agree closely for a converged basis
    s = O.SG4(2, 8)
    V = s.grid_hh(0.0)
    D = 2
    G = np.tile(np.eye(D).ravel(), s.sum_nq)
    one = np.ones(s.sum_nq)
    H1 = s.hpsi(np.eye(s.nb_ba), V, -0.5 * np.ones(D))
    H10 = s.hpsi(np.eye(s.nb_ba), V, None, type_op=10, G=G, Jac=one, sqRJ=one)
    e1 = np.linalg.eigvalsh(0.5 * (H1 + H1.T))[:3]
    e10 = np.linalg.eigvalsh(0.5 * (H10 + H10.T))[:3]
    assert np.allclose(e1, [1, 2, 2], atol=1e-10)
    assert np.allclose(e10, [1, 2, 2], atol=1e-8)


@pytest.mark.parametrize("nq", [1, 4, 11, 21])
def test_legendre_basis_orthonormal(nq):
    x, w, d0, d1, d2 = O.legendre_basis(nq, nq)
    assert np.max(np.abs(d0.T @ (w[:, None] * d0) - np.eye(nq))) < 1e-13
    assert np.all(np.diff(x) > 0) and x[0] > 0 and x[-1] < np.pi
    # P_l(cos t)'' (t) = -l(l+1) P_l - cot(t) P_l'  (Legendre ODE in theta)
    l = np.arange(nq)
    lhs = d2 + (np.cos(x) / np.sin(x))[:, None] * d1
    assert np.max(np.abs(lhs + l * (l + 1) * d0)) < 1e-9


def test_type10_legendre_rigid_rotor_spectrum():
    """-1/2 (1/sin t) d/dt sin t d/dt on the Legendre grid (type 10 with
    G = 1, Jac = sin t, sqRJ = 1): eigenvalues l(l+1)/2, exactly for the
    functions whose image stays in the basis (l <= nb - 2)."""
    L = 8
    s = O.SG4(1, L, kind=1)
    nq = s.sum_nq
    th = s.term_grid(0)[:, 0]
    H = s.hpsi(np.eye(s.nb_ba), np.zeros(nq), None, type_op=10, G=np.ones(nq),
               Jac=np.sin(th), sqRJ=np.ones(nq))
    nb = s.nb_ba
    l = np.arange(nb - 1)
    # the operator is diagonal on P_0..P_{nb-2} with l(l+1)/2 on the diagonal
    assert np.max(np.abs(np.diag(H)[: nb - 1] - l * (l + 1) / 2)) < 1e-9
    off = H[: nb - 1, : nb - 1] - np.diag(np.diag(H)[: nb - 1])
    assert np.max(np.abs(off)) < 1e-9


def test_type10_jacobi_3d_ground_level_bound():
    # the type-10 SG4 operator of the 3D Jacobi model is symmetric and its
    # lowest level lies above the sum of the two Morse harmonic ZPEs minus
    # anharmonic corrections (closed form for Morse: hw/2 - (hw)^2/(16 D))
    import elvibrot_amd.potentials as P
    s = O.SG4(3, 5, kind=[0, 0, 1], Q0=[3.0, 2.0, 0.0], nq_A=2, nb_A=2,
              scaleQ=[np.sqrt(P.AMU * np.sqrt(0.4 / P.AMU))] * 2 + [1.0])
    q = np.concatenate([s.term_grid(t) for t in range(s.nterms)])
    R, r, th = q.T
    V = P.morse(R, 0.2, 1.0, 3.0) + P.morse(r, 0.2, 1.0, 2.0) + 0.01 * np.cos(th) ** 2
    G = np.zeros((len(R), 3, 3))
    G[:, 0, 0] = G[:, 1, 1] = 1 / P.AMU
    G[:, 2, 2] = 1 / (P.AMU * R ** 2) + 1 / (P.AMU * r ** 2)
    H = s.hpsi(np.eye(s.nb_ba), V, None, type_op=10, G=G.reshape(-1), Jac=np.sin(th),
               sqRJ=np.ones_like(th))
    # d1 G d1 with pseudo-inverse derivative matrices is not exactly
    # Hermitian on a finite grid (the reference measures this,
    # sub_hermitic_H in sub_module_Davidson.f90:300-315); the lowest levels
    # are real and well separated
    w = np.sqrt(0.4 / P.AMU)
    e_morse = 2 * (w / 2 - w ** 2 / (16 * 0.2))
    ev = np.linalg.eigvals(H)
    e0 = np.min(ev.real)
    assert abs(ev[np.argmin(ev.real)].imag) < 1e-12
    assert e_morse < e0 < e_morse + 0.0101
    # converged in L: the same lowest level at L = 5 (above) and L = 8
    s8 = O.SG4(3, 8, kind=[0, 0, 1], Q0=[3.0, 2.0, 0.0], nq_A=2, nb_A=2,
               scaleQ=[np.sqrt(P.AMU * np.sqrt(0.4 / P.AMU))] * 2 + [1.0])
    q8 = np.concatenate([s8.term_grid(t) for t in range(s8.nterms)])
    R8, r8, t8 = q8.T
    V8 = P.morse(R8, 0.2, 1.0, 3.0) + P.morse(r8, 0.2, 1.0, 2.0) + 0.01 * np.cos(t8) ** 2
    G8 = np.zeros((len(R8), 3, 3))
    G8[:, 0, 0] = G8[:, 1, 1] = 1 / P.AMU
    G8[:, 2, 2] = 1 / (P.AMU * R8 ** 2) + 1 / (P.AMU * r8 ** 2)
    H8 = s8.hpsi(np.eye(s8.nb_ba), V8, None, type_op=10, G=G8.reshape(-1), Jac=np.sin(t8),
                 sqRJ=np.ones_like(t8))
    assert abs(np.min(np.linalg.eigvals(H8).real) - e0) < 1e-5 * e0


@pytest.mark.parametrize("D,L", [(3, 4), (6, 3), (6, 6)])
def test_extended_precision_action_matches(D, L):
    """orc_hpsi_ld is the same discrete operator as orc_hpsi evaluated in long
    double: the two agree within eps times the conditioning bound of the
    Smolyak sum, and the extended one is linear to far below FP64 rounding."""
    o = O.SG4(D, L)
    V = o.grid_hh(0.111803)
    ck = -0.5 * np.ones(D)
    rng = np.random.default_rng(D * 10 + L)
    psi = rng.standard_normal((o.nb_ba, 3))
    a, b = o.hpsi(psi, V, ck), o.hpsi_ld(psi, V, ck)
    bound = o.absbound(psi, V, ck)
    assert np.max(np.abs(a - b)) <= 2 * np.finfo(float).eps * np.max(bound)
    lin = o.hpsi_ld(2 * psi[:, :1] - 3 * psi[:, 1:2], V, ck)[:, 0]
    assert np.max(np.abs(lin - (2 * b[:, 0] - 3 * b[:, 1]))) < 1e-14 * np.max(np.abs(lin))


def _grid_points(o):
    return np.concatenate([o.term_grid(t) for t in range(o.nterms)])


def test_general_type1_reduces_to_constant_diagonal():
    """orc_hpsi_t1g with the identity term (F = V) and d2 terms of constant
    coefficient equals the type-1 action with constant c_k."""
    D, L = 4, 4
    o = O.SG4(D, L)
    V = o.grid_hh(0.111803)
    ck = np.array([-0.5, -0.3, -0.7, -0.5])
    psi = np.random.default_rng(1).standard_normal((o.nb_ba, 3))
    der = [(-1, -1)] + [(k, k) for k in range(D)]
    F = np.vstack([V] + [ck[k] * np.ones(o.sum_nq) for k in range(D)])
    a, b = o.hpsi(psi, V, ck), o.hpsi_t1g(psi, der, F)
    assert np.max(np.abs(a - b)) < 1e-14 * np.max(np.abs(a))


def test_general_type1_sheared_oscillator_spectrum():
    """Mixed second derivatives: H = -1/2 sum_ij G_ij d_i d_j + V with
    V = 1/2 |Omega A q|^2 and G = A^-1 A^-T is a harmonic oscillator in the
    sheared coordinates y = A q, spectrum sum_k w_k (n_k + 1/2) exactly; the
    2D L=12 Smolyak operator in q reproduces its 5 lowest levels to 1e-11
    (swapping a mode index of the mixed terms or dropping them fails)."""
    o = O.SG4(2, 12)
    q = _grid_points(o)
    A = np.array([[1.0, 0.3], [0.0, 1.0]])
    w = np.array([1.0, 1.3])
    V = 0.5 * np.sum((w ** 2) * (q @ A.T) ** 2, axis=1)
    Gm = np.linalg.inv(A) @ np.linalg.inv(A).T
    der = [(-1, -1), (0, 0), (1, 1), (0, 1), (1, 0)]
    one = np.ones(len(V))
    F = np.vstack([V, -0.5 * Gm[0, 0] * one, -0.5 * Gm[1, 1] * one, -0.5 * Gm[0, 1] * one,
                   -0.5 * Gm[1, 0] * one])
    ev = np.sort(np.linalg.eigvals(o.hpsi_t1g(np.eye(o.nb_ba), der, F)).real)[:5]
    ex = np.sort([w[0] * (a + 0.5) + w[1] * (b + 0.5) for a in range(6) for b in range(6)])[:5]
    assert np.max(np.abs(ev - ex)) < 1e-11


def test_general_type1_equals_type10_form_when_converged():
    """Grid-dependent coefficients and first derivatives: the operator
    -1/2 sum_ij d_i G_ij(q) d_j + V in its type-10 form (d1 G d1) and in its
    type-1 form (-1/2 G_ij d_i d_j - 1/2 (sum_i d_i G_ij) d_j, the derivative
    of G taken analytically) converge to the same levels; at 2D L=12 the 4
    lowest agree to 1e-10 (dropping the first-derivative terms moves them by
    ~1e-3)."""
    o = O.SG4(2, 12)
    q = _grid_points(o)
    n = len(q)
    g00, g11, g01 = 1 + 0.1 * q[:, 1] ** 2, 1 + 0.1 * q[:, 0] ** 2, 0.05 * q[:, 0] * q[:, 1]
    V = 0.5 * (q[:, 0] ** 2 + 1.2 * q[:, 1] ** 2)
    G = np.zeros((n, 2, 2))
    G[:, 0, 0], G[:, 1, 1], G[:, 0, 1], G[:, 1, 0] = g00, g11, g01, g01
    H10 = o.hpsi(np.eye(o.nb_ba), V, None, type_op=10, G=G.reshape(n, 4), Jac=np.ones(n),
                 sqRJ=np.ones(n))
    der = [(-1, -1), (0, 0), (1, 1), (0, 1), (1, 0), (0, -1), (1, -1)]
    F = np.vstack([V, -0.5 * g00, -0.5 * g11, -0.5 * g01, -0.5 * g01, -0.025 * q[:, 0],
                   -0.025 * q[:, 1]])
    e10 = np.sort(np.linalg.eigvals(H10).real)[:4]
    e1 = np.sort(np.linalg.eigvals(o.hpsi_t1g(np.eye(o.nb_ba), der, F)).real)[:4]
    assert np.max(np.abs(e1 - e10)) < 1e-10
    e1x = np.sort(np.linalg.eigvals(o.hpsi_t1g(np.eye(o.nb_ba), der[:5], F[:5])).real)[:4]
    assert np.max(np.abs(e1x - e10)) > 1e-4


def test_type10_nonsymmetric_G_jacobian_and_density_expanded():
    """Type 10 with every factor non-trivial (sub_OpPsi_SG4.f90:361-403):
    H psi = -1/2 (1/(J s)) sum_ij d_i [J G_ji d_j (s psi)] + V psi with a
    position-dependent, non-symmetric G (G_ji multiplies d_j inside d_i, the
    reference's GGiq(:,j,i)), Jac J != 1 and sqRhoOVERJac s != 1.  Expanded by
    the product rule it is a general type-1 operator,
      -1/2 G_ji d_i d_j
      - 1/(2 J s) [s sum_i d_i(J G_ki) + J sum_i (G_ki + G_ik) d_i s] d_k
      - 1/(2 J s) sum_ij [d_i(J G_ji) d_j s + J G_ji d_i d_j s] + V,
    with the derivatives of J, s and G taken analytically.  At 2D L=12 the two
    discrete operators give the same Galerkin block on the 10 lowest
    oscillator states to 1e-11; with G transposed (the antisymmetric part of G
    enters as a first-derivative drift), or with s left out of the type-10
    form, they differ by more than 1e-2."""
    o = O.SG4(2, 12)
    q = _grid_points(o)
    n, x, y = len(q), q[:, 0], q[:, 1]
    Hho = o.hpsi(np.eye(o.nb_ba), 0.5 * (x ** 2 + y ** 2), -0.5 * np.ones(2))
    U = np.linalg.eigh(0.5 * (Hho + Hho.T))[1][:, :10]
    J, dJ = 1 + 0.1 * x ** 2 + 0.05 * y ** 2, [0.2 * x, 0.1 * y]
    s, ds = 1 + 0.05 * x ** 2, [0.1 * x, 0 * x]
    dds = [[0.1 + 0 * x, 0 * x], [0 * x, 0 * x]]
    a = 0.1 * y + 0.05 * x
    Gs = [[1 + 0.1 * y ** 2, 0.05 * x * y + a], [0.05 * x * y - a, 1 + 0.1 * x ** 2]]
    # dG[i][j][k] = d_k G_ij
    dGs = [[[0 * x, 0.2 * y], [0.05 * y + 0.05, 0.05 * x + 0.1]],
           [[0.05 * y - 0.05, 0.05 * x - 0.1], [0.2 * x, 0 * x]]]
    V = 0.5 * (x ** 2 + 1.2 * y ** 2)

    def type10(G, sq):
        Gm = np.zeros((n, 2, 2))
        for i in range(2):
            for j in range(2):
                Gm[:, j, i] = G[i][j]          # G[i + D j + D^2 p] = G_ij
        return o.hpsi(np.eye(o.nb_ba), V, None, type_op=10, G=Gm.reshape(n, 4), Jac=J, sqRJ=sq)

    def expanded(G, dG):
        dJG = lambda i, j: dJ[i] * G[j][i] + J * dG[j][i][i]     # d_i (J G_ji)
        c0 = sum(dJG(i, j) * ds[j] + J * G[j][i] * dds[i][j] for i in range(2) for j in range(2))
        der, F = [(-1, -1)], [V - c0 / (2 * J * s)]
        for i in range(2):
            for j in range(2):
                der.append((i, j))
                F.append(-0.5 * G[j][i])
        for k in range(2):
            ck = sum(s * dJG(i, k) + J * (G[k][i] + G[i][k]) * ds[i] for i in range(2))
            der.append((k, -1))
            F.append(-ck / (2 * J * s))
        return o.hpsi_t1g(np.eye(o.nb_ba), der, np.vstack(F))

    A = U.T @ type10(Gs, s) @ U
    B = U.T @ expanded(Gs, dGs) @ U
    assert np.max(np.abs(A - B)) < 1e-11
    GT = [[Gs[j][i] for j in range(2)] for i in range(2)]
    dGT = [[dGs[j][i] for j in range(2)] for i in range(2)]
    assert np.max(np.abs(A - U.T @ expanded(GT, dGT) @ U)) > 1e-2
    assert np.max(np.abs(U.T @ type10(Gs, np.ones(n)) @ U - B)) > 1e-2


def _hh_dense(D, L):
    o = O.SG4(D, L)
    V = o.grid_hh(0.111803)
    mu = np.concatenate([o.term_grid(t) for t in range(o.nterms)]).sum(1)
    ck = -0.5 * np.ones(D)
    hfun = lambda X, e: o.hpsi(X, V + e * mu, ck)
    I = np.eye(o.nb_ba)
    return o, hfun, o.hpsi(I, V, ck), o.hpsi(I, mu, np.zeros(D))


def test_taylor_oracle_equals_exponential_under_a_field():
    """Taylor propagation (march_noD_field) with segment-constant fields equals
    prod_s exp(-i (H + eps_s M) dt)^steps on the dense matrices of the same
    operator; a wrong sign of i, a dropped 1/k or a missing E0 phase fails."""
    import scipy.linalg as sla
    o, hfun, H, M = _hh_dense(3, 3)
    ev = np.linalg.eigvalsh(0.5 * (H + H.T))
    E0, dt, eps = 0.5 * (ev[0] + ev[-1]), 0.1, [0.05, -0.03, 0.02]
    rng = np.random.default_rng(5)
    psi = rng.standard_normal((o.nb_ba, 2)) + 1j * rng.standard_normal((o.nb_ba, 2))
    got, used = O.taylor_propagate(hfun, psi, dt, E0, 400, 1e-28, eps, 2)
    ref = psi.copy()
    for e in eps:
        U = sla.expm(-1j * (H + e * M) * dt)
        ref = U @ (U @ ref)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
    assert used.shape == (6, 2) and used.min() > 5
    bad, _ = O.taylor_propagate(lambda X, e: -hfun(X, e), psi, dt, -E0, 400, 1e-28, eps, 2)
    assert np.max(np.abs(bad - ref)) > 1e-2


def test_taylor_oracle_cut_rule():
    """Terms used per step = first k with ||(-i dt)^k/k! (H - E0)^k psi||^2 below
    poly_tol, from the eigen-decomposition (per state, as the reference's loop
    over the packets; the block rule takes the largest)."""
    o, hfun, H, _ = _hh_dense(2, 4)
    Hs = 0.5 * (H + H.T)
    lam, U = np.linalg.eigh(Hs)
    E0, dt, tol = 0.5 * (lam[0] + lam[-1]), 0.2, 1e-20
    rng = np.random.default_rng(6)
    psi = rng.standard_normal((o.nb_ba, 3)) + 0j
    psi[:, 0] = U[:, np.argmin(np.abs(lam - 0.5 * (lam[0] + lam[-1])))]   # eigenvalue near E0: fastest
    sym = lambda X, e: Hs @ X
    c = U.T @ psi
    expect = []
    for s in range(3):
        k, term = 0, c[:, s].astype(complex)
        while True:
            k += 1
            term = term * (-1j * dt / k) * (lam - E0)
            if np.sum(np.abs(term) ** 2) < tol:
                break
        expect.append(k)
    _, per = O.taylor_propagate(sym, psi, dt, E0, 500, tol, per_state=True)
    _, blk = O.taylor_propagate(sym, psi, dt, E0, 500, tol)
    assert list(per[0]) == expect
    assert list(blk[0]) == [max(expect)] * 3
    assert expect[0] < expect[1]
