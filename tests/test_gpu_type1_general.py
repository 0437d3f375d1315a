"""General type-1 operator (grid-dependent coefficients, first, second and
mixed derivatives: sg4_plan_set_operator_type1) on the MI355X through the
C ABI: parity with the oracle at 1e-12 relative on random coefficient grids
covering every kind of derivative term (also with LB < LG), and the spectrum
of a sheared harmonic oscillator (mixed derivatives with constant
coefficients, exact levels sum_k w_k (n_k + 1/2)) from device applies."""
import numpy as np
import pytest

import elvibrot_amd as E
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DER = [(-1, -1), (0, -1), (1, 1), (0, 2), (2, 0), (1, -1), (2, 2), (0, 0), (-1, 2), (1, 2)]


@pytest.mark.parametrize("D,L,LB,nvec", [(3, 4, 4, 5), (3, 5, 3, 3), (4, 3, 3, 4)])
def test_general_type1_parity(cuda, D, L, LB, nvec):
    print(f"case D={D} L={L} LB={LB}", flush=True)
    b = E.SG4Basis(D, L, LB)
    o = O.SG4(D, L, LB)
    der = [d for d in DER if max(d) < D]
    rng = np.random.default_rng(D * 100 + L * 10 + LB)
    F = rng.standard_normal((len(der), o.sum_nq))
    op = E.SG4Operator(b, max_nvec=8, potential=("type1", der, F))
    psi = rng.standard_normal((b.nb_ba, nvec))
    got = op.apply_host(psi)
    ref = o.hpsi_t1g(psi, der, F)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"general type 1, {len(der)} terms: {err:.2e}")
    assert err < 1e-12


def test_general_type1_sheared_oscillator_on_device(cuda):
    import torch
    D, L = 2, 12
    b = E.SG4Basis(D, L)
    q = np.concatenate([b.term_grid(t) for t in range(b.nterms)])
    A = np.array([[1.0, 0.3], [0.0, 1.0]])
    w = np.array([1.0, 1.3])
    V = 0.5 * np.sum((w ** 2) * (q @ A.T) ** 2, axis=1)
    Gm = np.linalg.inv(A) @ np.linalg.inv(A).T
    der = [(-1, -1), (0, 0), (1, 1), (0, 1), (1, 0)]
    one = np.ones(len(V))
    F = np.vstack([V, -0.5 * Gm[0, 0] * one, -0.5 * Gm[1, 1] * one, -0.5 * Gm[0, 1] * one,
                   -0.5 * Gm[1, 0] * one])
    op = E.SG4Operator(b, max_nvec=16, potential=("type1", der, F))
    n = b.nb_ba
    H = np.zeros((n, n))
    for c0 in range(0, n, 16):
        c1 = min(c0 + 16, n)
        X = torch.zeros(c1 - c0, n, device=cuda, dtype=torch.float64)
        X[torch.arange(c1 - c0), torch.arange(c0, c1)] = 1.0
        Y = torch.empty_like(X)
        op.apply(X, Y)
        H[:, c0:c1] = Y.cpu().numpy().T
    ev = np.sort(np.linalg.eigvals(H).real)[:5]
    ex = np.sort([w[0] * (i + 0.5) + w[1] * (j + 0.5) for i in range(6) for j in range(6)])[:5]
    print("levels", ev, "exact", ex)
    assert np.max(np.abs(ev - ex)) < 1e-11


def test_general_type1_fp64_only(cuda):
    b = E.SG4Basis(3, 3)
    o = O.SG4(3, 3)
    op = E.SG4Operator(b, max_nvec=2, potential=("type1", [(-1, -1)], np.ones((1, o.sum_nq))))
    for prec in (E.PREC_F32, E.PREC_F32X2):
        with pytest.raises(E.SG4Error):
            op.apply_host(np.ones((b.nb_ba, 1)), prec)
