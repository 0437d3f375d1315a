"""Two-surface operators (nb0 = 2) composed from single-surface plans
(elvibrot_amd.surfaces), against closed forms and the oracle:
  * a constant coupling c between two copies of the Henon-Heiles surface is
    c times the identity on the packed basis (the Smolyak sum reproduces the
    sparse space exactly), so H = [[H0, c], [c, H0]] has the levels
    E0 - c and E0 + c of H0, and a packet started on surface 1 oscillates:
    psi(t) = (cos(ct) U psi0, -i sin(ct) U psi0), U = exp(-i H0 t);
  * a linear vibronic coupling kappa q_1 with a shifted second surface:
    every block against the oracle (1e-12), the symmetry of the operator,
    and the Davidson levels against the dense oracle matrix."""
import numpy as np
import pytest
import scipy.linalg as sla

import elvibrot_amd as E
from elvibrot_amd.surfaces import SG4Surfaces
from elvibrot_amd.davidson import sub_propagation_Davidson
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LAM = 0.111803


def _blocks(D, L, c):
    b = E.SG4Basis(D, L)
    H0 = E.SG4Operator(b, max_nvec=16)
    Cp = E.SG4Operator(b, max_nvec=16, potential=c * np.ones(b.sum_nq), ck=np.zeros(D))
    return b, H0, Cp


def test_constant_coupling_apply_and_levels(cuda):
    import torch
    D, L, c = 4, 4, 0.05
    b, H0, Cp = _blocks(D, L, c)
    S = SG4Surfaces([[H0, Cp], [Cp, H0]])
    o = O.SG4(D, L)
    V, ck = o.grid_hh(LAM), -0.5 * np.ones(D)
    nb = b.nb_ba
    rng = np.random.default_rng(31)
    psi = rng.standard_normal((2 * nb, 3))
    got = S.apply_host(psi)
    ref = np.vstack([o.hpsi(psi[:nb], V, ck) + c * psi[nb:], o.hpsi(psi[nb:], V, ck) + c * psi[:nb]])
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"2-surface apply vs H0 (+) c I: {err:.2e}")
    assert err < 1e-12
    Hd = o.hpsi(np.eye(nb), V, ck)
    lam = np.linalg.eigvalsh(0.5 * (Hd + Hd.T))
    expect = np.sort(np.concatenate([lam - c, lam + c]))[:8]
    r = sub_propagation_Davidson(S, 8, max_it=300, conv_resi=1e-8, conv_ene=1e-12)
    print("levels", r.E - expect)
    assert r.converged.all() and np.max(np.abs(r.E - expect)) < 1e-9


def test_constant_coupling_rabi_propagation(cuda):
    import torch
    from elvibrot_amd.propagation import ChebyshevPropagator, spectral_bounds
    D, L, c, dt, steps = 3, 4, 0.2, 0.1, 6
    b, H0, Cp = _blocks(D, L, c)
    S = SG4Surfaces([[H0, Cp], [Cp, H0]])
    o = O.SG4(D, L)
    nb = b.nb_ba
    Hd = o.hpsi(np.eye(nb), o.grid_hh(LAM), -0.5 * np.ones(D))
    Hd = 0.5 * (Hd + Hd.T)
    lo, hi = spectral_bounds(S)
    prop = ChebyshevPropagator(S, dt, lo, hi)
    rng = np.random.default_rng(32)
    p0 = rng.standard_normal(nb) + 1j * rng.standard_normal(nb)
    p0 /= np.linalg.norm(p0)
    x = torch.zeros(1, 2 * nb, dtype=torch.complex128, device=cuda)
    x[0, :nb] = torch.tensor(p0, device=cuda)
    for _ in range(steps):
        x, _ = prop.step(x)
    t = dt * steps
    u = sla.expm(-1j * Hd * t) @ p0
    ref = np.concatenate([np.cos(c * t) * u, -1j * np.sin(c * t) * u])
    err = np.max(np.abs(x[0].cpu().numpy() - ref))
    print(f"Rabi oscillation after t = {t}: {err:.2e}")
    assert err < 1e-10


def test_vibronic_coupling_against_oracle(cuda):
    import torch
    D, L, kappa, delta = 3, 5, 0.08, 0.3
    b = E.SG4Basis(D, L)
    o = O.SG4(D, L)
    q1 = np.concatenate([o.term_grid(t)[:, 0] for t in range(o.nterms)])
    V, ck = o.grid_hh(LAM), -0.5 * np.ones(D)
    H1 = E.SG4Operator(b, max_nvec=16, potential=V, ck=ck)
    H2 = E.SG4Operator(b, max_nvec=16, potential=V + delta, ck=ck)
    K = E.SG4Operator(b, max_nvec=16, potential=kappa * q1, ck=np.zeros(D))
    S = SG4Surfaces([[H1, K], [K, H2]])
    nb = b.nb_ba
    rng = np.random.default_rng(33)
    psi = rng.standard_normal((2 * nb, 5))
    blk = lambda x, W, c: o.hpsi(x, W, c)
    ref = np.vstack([blk(psi[:nb], V, ck) + blk(psi[nb:], kappa * q1, np.zeros(D)),
                     blk(psi[nb:], V + delta, ck) + blk(psi[:nb], kappa * q1, np.zeros(D))])
    got = S.apply_host(psi)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"vibronic 2-surface apply vs oracle blocks: {err:.2e}")
    assert err < 1e-12
    phi = rng.standard_normal(2 * nb)
    sym = abs(phi @ S.apply_host(psi[:, 0]) - psi[:, 0] @ S.apply_host(phi))
    assert sym < 1e-12 * np.linalg.norm(phi) * np.linalg.norm(psi[:, 0]) * 10
    M = np.block([[blk(np.eye(nb), V, ck), blk(np.eye(nb), kappa * q1, np.zeros(D))],
                  [blk(np.eye(nb), kappa * q1, np.zeros(D)), blk(np.eye(nb), V + delta, ck)]])
    expect = np.linalg.eigvalsh(0.5 * (M + M.T))[:6]
    r = sub_propagation_Davidson(S, 6, max_it=300, conv_resi=1e-8, conv_ene=1e-12)
    assert r.converged.all() and np.max(np.abs(r.E - expect)) < 1e-9
