"""Chebyshev propagation on the device:
  * one step against exp(-i H dt) of the dense FP64 oracle matrix (1e-10),
  * the norm is conserved and the adjoint step inverts the forward step,
  * field-driven propagation (piecewise-constant eps, dipole mu = sum q_i):
    dJ/deps (exact discrete adjoint, checkpointed recomputation) against
    central finite differences of J (1e-6 relative)."""
import numpy as np
import pytest
import scipy.linalg as sla

import elvibrot_amd as E
from elvibrot_amd.propagation import (ChebyshevPropagator, spectral_bounds, propagate_field,
                                      overlap_gradient)
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def hh(D, L):
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=8)
    return b, op


def test_chebyshev_step_vs_dense_exponential(cuda):
    import torch
    D, L, dt = 3, 4, 0.1
    b, op = hh(D, L)
    o = O.SG4(D, L)
    H = o.hpsi(np.eye(o.nb_ba), o.grid_hh(0.111803), -0.5 * np.ones(D))
    H = 0.5 * (H + H.T)
    lo, hi = spectral_bounds(op)
    ev = np.linalg.eigvalsh(H)
    assert lo <= ev[0] and hi >= ev[-1]
    prop = ChebyshevPropagator(op, dt, lo, hi)
    rng = np.random.default_rng(3)
    psi = rng.standard_normal((2, b.nb_ba)) + 1j * rng.standard_normal((2, b.nb_ba))
    psi /= np.linalg.norm(psi, axis=1)[:, None]
    x = torch.tensor(psi, device=cuda)
    y, n = prop.step(x)
    ref = (sla.expm(-1j * H * dt) @ psi.T).T
    assert np.max(np.abs(y.cpu().numpy() - ref)) < 1e-10
    z, _ = prop.step(y, conj=True, npoly=n)
    assert np.max(np.abs(z.cpu().numpy() - psi)) < 1e-10
    assert abs(np.linalg.norm(y.cpu().numpy()[0]) - 1) < 1e-12


def test_field_gradient_vs_finite_differences(cuda):
    import torch
    D, L = 3, 3
    b, op = hh(D, L)
    o = O.SG4(D, L)
    q = np.concatenate([o.term_grid(t) for t in range(o.nterms)])
    mu = torch.tensor(q.sum(1), device=cuda)
    nseg, steps, dt = 12, 5, 0.1
    eps = np.random.default_rng(7).normal(0, 0.01, nseg)
    lo, hi = spectral_bounds(op)
    w = 0.05 * (hi - lo) + np.abs(eps).max() * np.abs(q.sum(1)).max()
    prop = ChebyshevPropagator(op, dt, lo - w, hi + w, mu=mu)
    psi0 = torch.zeros(1, b.nb_ba, dtype=torch.complex128, device=cuda)
    psi0[0, 0] = 1.0
    target = torch.zeros_like(psi0)
    target[0, 1] = np.sqrt(0.5)
    target[0, 0] = np.sqrt(0.5)
    J, g = overlap_gradient(prop, psi0, target, eps, steps, checkpoint_every=4)
    _, _, npolys = propagate_field(prop, psi0, eps, steps)

    def Jof(e):   # the same discrete map: polynomial orders of the forward run
        psiT, _, _ = propagate_field(prop, psi0, e, steps, npolys=npolys)
        return abs(torch.sum(target.conj() * psiT).item()) ** 2

    assert abs(Jof(eps) - J) < 1e-13
    h = 1e-5
    for s in (0, 5, nseg - 1):
        ep, em = eps.copy(), eps.copy()
        ep[s] += h
        em[s] -= h
        fd = (Jof(ep) - Jof(em)) / (2 * h)
        print(f"segment {s}: adjoint {g[s]:.10e} fd {fd:.10e}")
        assert abs(g[s] - fd) < 1e-6 * max(abs(fd), 1e-6)


def test_batch_of_uneven_instances_equals_single_plans(cuda):
    """One batched plan over instances of different D and L: each instance's
    block of H.psi equals the oracle's (1e-12)."""
    specs = [(3, 2), (4, 5), (6, 3), (5, 4), (2, 8)]
    bases = [E.SG4Basis(D, L) for D, L in specs]
    op = E.SG4Batch(bases, max_nvec=4)
    rng = np.random.default_rng(8)
    psi = rng.standard_normal((op.basis.nb_ba, 3))
    got = op.apply_host(psi)
    for (D, L), r0, r1 in zip(specs, op.row0[:-1], op.row0[1:]):
        o = O.SG4(D, L)
        ref = o.hpsi(psi[r0:r1], o.grid_hh(0.111803), -0.5 * np.ones(D))
        assert np.max(np.abs(got[r0:r1] - ref)) / np.max(np.abs(ref)) < 1e-12


def test_batched_field_gradient_vs_finite_differences(cuda):
    """Batched propagation (3 uneven instances in one plan, one launch per
    H.psi) under a shared piecewise-constant field; J = sum_i |<t_i|psi_i(T)>|^2."""
    import torch
    specs = [(2, 3), (3, 2), (4, 2)]
    bases = [E.SG4Basis(D, L) for D, L in specs]
    op = E.SG4Batch(bases, max_nvec=4)
    q = np.concatenate([b.term_grid(t).sum(1) for b in bases for t in range(b.nterms)])
    mu = torch.tensor(q, device=cuda)
    nseg, steps, dt = 8, 4, 0.1
    eps = np.random.default_rng(9).normal(0, 0.02, nseg)
    lo, hi = spectral_bounds(op)
    w = 0.05 * (hi - lo) + np.abs(eps).max() * np.abs(q).max()
    prop = ChebyshevPropagator(op, dt, lo - w, hi + w, mu=mu)
    nb = op.basis.nb_ba
    psi0 = torch.zeros(1, nb, dtype=torch.complex128, device=cuda)
    target = torch.zeros_like(psi0)
    blocks = list(zip(op.row0[:-1], op.row0[1:]))
    for r0, r1 in blocks:
        psi0[0, r0] = 1.0
        target[0, r0] = np.sqrt(0.5)
        target[0, r0 + 1] = np.sqrt(0.5)
    J, g = overlap_gradient(prop, psi0, target, eps, steps, checkpoint_every=3, blocks=blocks)
    _, _, npolys = propagate_field(prop, psi0, eps, steps)

    def Jof(e):
        psiT, _, _ = propagate_field(prop, psi0, e, steps, npolys=npolys)
        return sum(abs(torch.sum(target[0, r0:r1].conj() * psiT[0, r0:r1]).item()) ** 2
                   for r0, r1 in blocks)

    assert abs(Jof(eps) - J) < 1e-13
    h = 1e-5
    for s in (0, nseg // 2, nseg - 1):
        ep, em = eps.copy(), eps.copy()
        ep[s] += h
        em[s] -= h
        fd = (Jof(ep) - Jof(em)) / (2 * h)
        print(f"segment {s}: adjoint {g[s]:.10e} fd {fd:.10e}")
        assert abs(g[s] - fd) < 1e-6 * max(abs(fd), 1e-6)


def small_batch(cuda, max_nvec=8):
    import torch
    specs = [(2, 3), (3, 2), (4, 2)]
    bases = [E.SG4Basis(D, L) for D, L in specs]
    op = E.SG4Batch(bases, max_nvec=max_nvec)
    q = np.concatenate([b.term_grid(t).sum(1) for b in bases for t in range(b.nterms)])
    return bases, op, q


def test_device_dipole_grid(cuda):
    bases, op, q = small_batch(cuda)
    mu = op.grid_dipole().cpu().numpy()
    assert np.max(np.abs(mu - q)) < 1e-14


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_F32X2])
def test_device_propagator_and_gradient(cuda, prec):
    """sg4_cheby_propagate / sg4_cheby_field_gradient (recurrence in device
    kernels, adjoint and checkpoints in the library) against the host-driven
    recurrence with the same fixed order count (1e-12 in FP64), and the
    device gradient against central finite differences of the device
    propagation (1e-6 relative in FP64; the split-precision mode against its
    FP64 gradient)."""
    import torch
    from elvibrot_amd.propagation import propagate_device, field_gradient_device
    bases, op, q = small_batch(cuda)
    mu = op.grid_dipole()
    nseg, steps, dt = 6, 3, 0.1
    eps = np.random.default_rng(9).normal(0, 0.02, nseg)
    lo, hi = spectral_bounds(op)
    w = 0.05 * (hi - lo) + np.abs(eps).max() * np.abs(q).max()
    prop = ChebyshevPropagator(op, dt, lo - w, hi + w, mu=mu, prec=prec)
    nb = op.basis.nb_ba
    rng = np.random.default_rng(4)
    psi0 = torch.tensor(rng.standard_normal((2, nb)) + 1j * rng.standard_normal((2, nb)), device=cuda)
    psi0 /= torch.linalg.norm(psi0, dim=1, keepdim=True)
    target = torch.zeros_like(psi0)
    blocks = list(zip(op.row0[:-1], op.row0[1:]))
    for r0, r1 in blocks:
        target[:, r0] = np.sqrt(0.5)
        target[:, r0 + 1] = np.sqrt(0.5)
    psiT = propagate_device(prop, psi0, eps, steps)
    if prec == E.PREC_FP64:
        ref, _, _ = propagate_field(prop, psi0, eps, steps,
                                    npolys=[[prop.nmax] * steps for _ in range(nseg)])
        err = torch.max(torch.abs(psiT - ref)).item()
        print(f"device vs host recurrence: {err:.2e}")
        assert err < 1e-12
    J, g = field_gradient_device(prop, psi0, target, eps, steps, checkpoint_every=4, blocks=blocks)

    def Jof(e):
        pT = propagate_device(prop, psi0, e, steps)
        return sum(abs(torch.sum(target[:, r0:r1].conj() * pT[:, r0:r1]).item()) ** 2
                   for r0, r1 in blocks)

    assert abs(Jof(eps) - J) < 1e-12 * max(1.0, J)
    if prec == E.PREC_FP64:
        h = 1e-5
        for s in (0, nseg // 2, nseg - 1):
            ep, em = eps.copy(), eps.copy()
            ep[s] += h
            em[s] -= h
            fd = (Jof(ep) - Jof(em)) / (2 * h)
            print(f"segment {s}: device adjoint {g[s]:.10e} fd {fd:.10e}")
            assert abs(g[s] - fd) < 1e-6 * max(abs(fd), 1e-6)
    else:
        p64 = ChebyshevPropagator(op, dt, lo - w, hi + w, mu=mu)
        J64, g64 = field_gradient_device(p64, psi0, target, eps, steps, checkpoint_every=4,
                                         blocks=blocks)
        rel = np.max(np.abs(g - g64)) / np.max(np.abs(g64))
        print(f"split-precision gradient vs FP64: {rel:.2e}, J {J:.12f} vs {J64:.12f}")
        assert rel < 1e-4 and abs(J - J64) < 1e-5 * J64


def _taylor_case(cuda):
    import torch
    D, L = 3, 3
    b, op = hh(D, L)
    o = O.SG4(D, L)
    q = np.concatenate([o.term_grid(t) for t in range(o.nterms)]).sum(1)
    lo, hi = spectral_bounds(op)
    w = 0.05 * (hi - lo) + 0.05 * np.abs(q).max()
    rng = np.random.default_rng(11)
    psi = rng.standard_normal((2, b.nb_ba)) + 1j * rng.standard_normal((2, b.nb_ba))
    psi /= np.linalg.norm(psi, axis=1)[:, None]
    return D, b, op, o, q, torch.tensor(q, device=cuda), lo - w, hi + w, psi


def test_taylor_device_vs_oracle_exponential_and_chebyshev(cuda):
    """sg4_taylor_propagate (march_noD_field, field constant per segment)
    against the oracle's Taylor propagation with the same block cut (same
    terms per step, 1e-12), the dense exponentials of the oracle matrices
    (1e-10) and the device Chebyshev propagation of the same field (1e-9)."""
    import torch
    from elvibrot_amd.propagation import propagate_taylor_device, propagate_device, taylor_npoly
    D, b, op, o, q, mu, Hmin, Hmax, psi = _taylor_case(cuda)
    eps, steps, dt = [0.05, -0.03, 0.02], 2, 0.1
    got, orders = propagate_taylor_device(op, torch.tensor(psi, device=cuda), dt, Hmin, Hmax, eps,
                                          steps, mu=mu)
    got = got.cpu().numpy()
    E0 = Hmin + 0.5 * (Hmax - Hmin)
    V, ck = o.grid_hh(0.111803), -0.5 * np.ones(D)
    npoly = taylor_npoly(0.5 * (Hmax - Hmin) * dt)
    ref, used = O.taylor_propagate(lambda X, e: o.hpsi(X, V + e * q, ck), psi.T, dt, E0, npoly,
                                   1e-20, eps, steps)
    print(f"terms per step: device {orders.tolist()}, oracle {used[:, 0].tolist()}")
    assert orders.tolist() == used[:, 0].tolist()
    assert np.max(np.abs(got - ref.T)) < 1e-12
    H = o.hpsi(np.eye(o.nb_ba), V, ck)
    M = o.hpsi(np.eye(o.nb_ba), q, np.zeros(D))
    ex = psi.T.copy()
    for e in eps:
        U = sla.expm(-1j * (H + e * M) * dt)
        for _ in range(steps):
            ex = U @ ex
    assert np.max(np.abs(got - ex.T)) < 1e-10
    assert np.max(np.abs(np.linalg.norm(got, axis=1) - 1)) < 1e-10
    prop = ChebyshevPropagator(op, dt, Hmin, Hmax, mu=mu)
    ch = propagate_device(prop, torch.tensor(psi, device=cuda), eps, steps).cpu().numpy()
    assert np.max(np.abs(got - ch)) < 1e-9


def test_taylor_device_errors_restore_the_field(cuda):
    """The reference's STOPs become SG4_ERR_STATE: series not cut within npoly
    terms, a term's squared norm above 1e15 (time step too large), a state's
    squared norm above max_norm2; the field-free operator is back afterwards
    (to the rounding of the FP64 scatter-adds, whose order varies)."""
    import torch
    from elvibrot_amd.propagation import propagate_taylor_device
    D, b, op, o, q, mu, Hmin, Hmax, psi = _taylor_case(cuda)
    x = np.random.default_rng(2).standard_normal((b.nb_ba, 2))
    h0 = op.apply_host(x)
    t = torch.tensor(psi, device=cuda)
    same = lambda a, b: np.max(np.abs(a - b)) < 1e-13 * np.max(np.abs(b))
    with pytest.raises(E.SG4Error, match="poly_tol"):
        propagate_taylor_device(op, t, 0.1, Hmin, Hmax, [0.3], 1, mu=mu, npoly=3)
    assert same(op.apply_host(x), h0)
    with pytest.raises(E.SG4Error, match="reduce the time step"):
        propagate_taylor_device(op, t, 20.0, Hmin, Hmax, [0.3], 1, mu=mu, npoly=400)
    assert same(op.apply_host(x), h0)
    with pytest.raises(E.SG4Error, match="max_norm2"):
        propagate_taylor_device(op, t, 0.1, Hmin, Hmax, [0.3], 1, mu=mu, max_norm2=0.5)
    assert same(op.apply_host(x), h0)
    _, orders = propagate_taylor_device(op, t, 0.1, Hmin, Hmax, [0.3], 1, mu=mu, max_norm2=1.3)
    assert orders[0] > 5
