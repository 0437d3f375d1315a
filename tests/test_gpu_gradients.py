"""Eigenvalue gradients with respect to potential grid values and potential
parameters (device kernel sg4_grad_V), checked against
  * the exact linearity identity  sum_p g(p) U(p) = <phi| H[V=U, T=0] |psi>
    with the oracle's operator (1e-12),
  * finite differences of the eigenvalues (1e-6 relative),
  * the Hellmann-Feynman expectation value <psi|dV/dtheta|psi> computed by
    the oracle's operator with V = dV/dtheta and no kinetic term."""
import numpy as np
import pytest

import elvibrot_amd as E
import elvibrot_amd.potentials as P
from elvibrot_amd.gradients import grad_V, eigenvalue_gradients, dense_matrix, lowest_eigenpairs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_gradV_linearity_identity(cuda):
    D, L = 5, 4
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=4)
    o = O.SG4(D, L)
    rng = np.random.default_rng(5)
    phi = rng.standard_normal((b.nb_ba, 3))
    psi = rng.standard_normal((b.nb_ba, 3))
    g = grad_V(op, phi, psi)
    U = rng.standard_normal(o.sum_nq)
    lhs = g.T @ U
    rhs = np.einsum("iv,iv->v", phi, o.hpsi(psi, U, np.zeros(D)))
    assert np.max(np.abs(lhs - rhs)) / np.max(np.abs(rhs)) < 1e-12


def test_gradV_type10_linearity_identity(cuda):
    op = P.jacobi_triatomic_operator(5, max_nvec=4)
    b = op.basis
    sc = np.sqrt(P.AMU * np.sqrt(0.4 / P.AMU))
    o = O.SG4(3, 5, kind=[0, 0, 1], Q0=[3.0, 2.0, 0.0], scaleQ=[sc, sc, 1.0], nq_A=2, nb_A=2)
    rng = np.random.default_rng(6)
    phi = rng.standard_normal((b.nb_ba, 2))
    psi = rng.standard_normal((b.nb_ba, 2))
    g = grad_V(op, phi, psi)
    U = rng.standard_normal(o.sum_nq)
    n = o.sum_nq
    rhs = np.einsum("iv,iv->v", phi, o.hpsi(psi, U, None, type_op=10, G=np.zeros(9 * n),
                                               Jac=np.ones(n), sqRJ=np.ones(n)))
    assert np.max(np.abs(g.T @ U - rhs)) / np.max(np.abs(rhs)) < 1e-12


def test_henon_heiles_dE_dlambda_vs_finite_difference(cuda):
    D, L, n, lam, h = 4, 5, 5, 0.111803, 1e-4
    b = E.SG4Basis(D, L)
    o = O.SG4(D, L)
    q = np.concatenate([o.term_grid(t) for t in range(o.nterms)])
    dV = (q[:, :-1] ** 2 * q[:, 1:] - q[:, 1:] ** 3 / 3).sum(1)
    Vl = lambda x: 0.5 * (q ** 2).sum(1) + x * dV
    op = E.SG4Operator(b, max_nvec=8, potential=Vl(lam), ck=-0.5 * np.ones(D))
    Ev, dE, _ = eigenvalue_gradients(op, [dV], n, method="dense")
    fd = []
    for s in (+1, -1):
        ops = E.SG4Operator(b, max_nvec=8, potential=Vl(lam + s * h), ck=-0.5 * np.ones(D))
        fd.append(np.linalg.eigvalsh(dense_matrix(ops))[:n])
    fd = (fd[0] - fd[1]) / (2 * h)
    # Hellmann-Feynman with the oracle operator
    H = o.hpsi(np.eye(o.nb_ba), Vl(lam), -0.5 * np.ones(D))
    w, X = np.linalg.eigh(0.5 * (H + H.T))
    hf = np.einsum("iv,iv->v", X[:, :n], o.hpsi(X[:, :n], dV, np.zeros(D)))
    print("dE/dlambda", dE[:, 0], "fd", fd, "hf", hf)
    assert np.max(np.abs(dE[:, 0] - hf)) < 1e-10 * np.max(np.abs(hf)) + 1e-12
    assert np.max(np.abs(dE[:, 0] - fd) / np.maximum(np.abs(fd), 1e-3)) < 1e-6


def test_jacobi_5_levels_parameter_gradients_vs_finite_difference(cuda):
    """Config 'differentiation, curvilinear': 5 lowest levels of the 3D Jacobi
    triatomic (type 10), gradients w.r.t. Morse D, Morse a and the cos^2
    coupling, against central differences (1e-6 relative)."""
    L = 6
    base = dict(D=0.2, a=1.0, coupling=0.01)
    b = P.jacobi_triatomic_basis(L)
    q = np.concatenate([b.term_grid(t) for t in range(b.nterms)])
    R, r, th = q.T

    def grids(D, a, coupling):
        return P.jacobi_triatomic_grids(b, D=D, a=a, coupling=coupling)

    def levels(**kw):
        V, G, Jac, sqRJ = grids(**kw)
        op = E.SG4Operator(b, max_nvec=8, potential=V, type_op=10, G=G, Jac=Jac, sqRJ=sqRJ)
        return op

    mor = lambda x, D, a, xe: (1 - np.exp(-a * (x - xe)))
    dV_dD = mor(R, 0, 1.0, 3.0) ** 2 + mor(r, 0, 1.0, 2.0) ** 2
    dV_da = (2 * 0.2 * mor(R, 0, 1.0, 3.0) * np.exp(-(R - 3.0)) * (R - 3.0)
             + 2 * 0.2 * mor(r, 0, 1.0, 2.0) * np.exp(-(r - 2.0)) * (r - 2.0))
    dV_dc = np.cos(th) ** 2
    op = levels(**base)
    Ev, dE, _ = eigenvalue_gradients(op, [dV_dD, dV_da, dV_dc], 5)
    # levels 1 and 2 (R and r stretch fundamentals) are nearly degenerate
    # (identical Morse oscillators and masses), so the central difference
    # needs a small step: its error grows like h^2 / gap^2
    for j, key, h in ((0, "D", 1e-7), (1, "a", 1e-7), (2, "coupling", 1e-7)):
        ev = []
        for s in (+1, -1):
            kw = dict(base)
            kw[key] = base[key] + s * h
            Hs = dense_matrix(levels(**kw))
            ev.append(np.sort(np.linalg.eigvals(Hs).real)[:5])
        fd = (ev[0] - ev[1]) / (2 * h)
        print(key, dE[:, j], fd)
        assert np.max(np.abs(dE[:, j] - fd) / np.maximum(np.abs(fd), 1e-4)) < 1e-6


@pytest.mark.timeout(600)
def test_jacobi_L10_gradients_hellmann_feynman_and_finite_differences(cuda):
    """The curvilinear differentiation workload at its size: 3D Jacobi
    triatomic (type 10, grids evaluated on the device), Smolyak L=10 (2288
    packed rows).  Gradients of the 5 lowest levels with respect to Morse D,
    Morse a, the cos^2 coupling and one random direction of the grid values,
    from the device grid-value gradient contracted with dV/dtheta, against
      * Hellmann-Feynman <l|dH/dtheta|r>/<l|r>, dH/dtheta applied as an SG4
        operator of its own (V = dV/dtheta, G = 0): 1e-10 relative;
      * central differences of the levels, the grids re-evaluated on the
        device at theta +- h on the same basis: 1e-6 relative."""
    import time
    t0 = time.time()
    L, nlev = 10, 5
    base = dict(D=0.2, a=1.0, coupling=0.01)
    op = P.jacobi_triatomic_operator(L, max_nvec=64, **base)
    b = op.basis
    print(f"L={L} nb_ba={b.nb_ba} sum_nq={b.sum_nq}", flush=True)
    q = np.concatenate([b.term_grid(t) for t in range(b.nterms)])
    R, r, th = q.T
    ex = lambda x, xe: np.exp(-(x - xe))
    dV = {"D": (1 - ex(R, 3.0)) ** 2 + (1 - ex(r, 2.0)) ** 2,
          "a": (2 * 0.2 * (1 - ex(R, 3.0)) * ex(R, 3.0) * (R - 3.0)
                + 2 * 0.2 * (1 - ex(r, 2.0)) * ex(r, 2.0) * (r - 2.0)),
          "coupling": np.cos(th) ** 2,
          "grid": np.random.default_rng(41).uniform(0.0, 1.0, b.sum_nq)}
    keys = list(dV)
    # eigenvalue_gradients, with the eigenvectors kept for the check below
    Ev, Rv, Lv = lowest_eigenpairs(dense_matrix(op), nlev)
    g = grad_V(op, Lv, Rv)
    dE = np.stack([g.T @ dV[k] for k in keys], axis=1)
    print(f"levels {Ev} ({time.time() - t0:.1f} s)", flush=True)
    # Hellmann-Feynman with the derivative operator built on the product path
    _, _, Jac, sqRJ = P.jacobi_triatomic_grids(b, **base)
    for j, k in enumerate(keys):
        opd = E.SG4Operator(b, 64, potential=dV[k], type_op=10, G=np.zeros(9 * b.sum_nq),
                            Jac=Jac, sqRJ=sqRJ)
        hf = np.einsum("iv,iv->v", Lv, dense_matrix(opd) @ Rv)
        print(k, "grad", dE[:, j], "HF", hf, flush=True)
        assert np.max(np.abs(dE[:, j] - hf)) < 1e-10 * np.max(np.abs(hf))
    # central differences (levels 1 and 2 are nearly degenerate: small steps)
    V0, G0, _, _ = P.jacobi_triatomic_grids(b, **base)
    for j, k in enumerate(keys):
        h = 1e-7
        ev = []
        for s in (+1, -1):
            if k == "grid":
                o2 = E.SG4Operator(b, 64, potential=V0 + s * h * dV[k], type_op=10, G=G0,
                                   Jac=Jac, sqRJ=sqRJ)
            else:
                # same basis (its HO scaling is set by D and a), grids
                # re-evaluated on the device at the shifted parameter
                kw = dict(base)
                kw[k] = base[k] + s * h
                o2 = E.SG4Operator(b, 64, potential=("jacobi_triatomic", (
                    P.AMU, P.AMU, kw["D"], kw["a"], 3.0, 2.0, kw["coupling"])))
            ev.append(np.sort(np.linalg.eigvals(dense_matrix(o2)).real)[:nlev])
        fd = (ev[0] - ev[1]) / (2 * h)
        print(k, "fd", fd, f"({time.time() - t0:.1f} s)", flush=True)
        assert np.max(np.abs(dE[:, j] - fd) / np.maximum(np.abs(fd), 1e-4)) < 1e-6


def test_davidson_gradients_equal_dense_and_scale(cuda):
    """Eigenvalue gradients from the device Davidson's eigenvectors (no dense
    matrix) equal the dense path's at 5D L=5 (1e-8), and at 8D L=6 (nb_ba
    beyond a dense host diagonalisation) the ground state's dE/dlambda agrees
    with central differences of Davidson ground energies (1e-6 relative)."""
    def hh(D, L, lam):
        b = E.SG4Basis(D, L)
        q = np.concatenate([b.term_grid(t) for t in range(b.nterms)])
        dV = (q[:, :-1] ** 2 * q[:, 1:] - q[:, 1:] ** 3 / 3).sum(1)
        V = lambda x: 0.5 * (q ** 2).sum(1) + x * dV
        return b, dV, V
    lam = 0.111803
    b, dV, V = hh(5, 5, lam)
    op = E.SG4Operator(b, max_nvec=16, potential=V(lam), ck=-0.5 * np.ones(5))
    Ed, gd, _ = eigenvalue_gradients(op, [dV], 3, method="dense")
    Ev, gv, _ = eigenvalue_gradients(op, [dV], 3, method="davidson")
    print("5D L=5:", Ev - Ed, gv[:, 0] - gd[:, 0])
    assert np.max(np.abs(Ev - Ed)) < 1e-9 and np.max(np.abs(gv - gd)) < 1e-8
    from elvibrot_amd.davidson import sub_propagation_Davidson
    b, dV, V = hh(8, 6, lam)
    print(f"8D L=6: nb_ba {b.nb_ba}", flush=True)
    op = E.SG4Operator(b, max_nvec=16, potential=V(lam), ck=-0.5 * np.ones(8))
    Ev, gv, _ = eigenvalue_gradients(op, [dV], 1)
    h, e = 1e-4, []       # central-difference truncation ~ h^2 |d3E/dlambda3| / 6 ~ 2e-8
    for s in (+1, -1):
        o2 = E.SG4Operator(b, max_nvec=16, potential=V(lam + s * h), ck=-0.5 * np.ones(8))
        e.append(sub_propagation_Davidson(o2, 1, max_it=500, conv_resi=1e-9, conv_ene=1e-13).E[0])
    fd = (e[0] - e[1]) / (2 * h)
    print(f"dE0/dlambda {gv[0, 0]:.10e} fd {fd:.10e}")
    assert abs(gv[0, 0] - fd) < 1e-6 * abs(fd)
