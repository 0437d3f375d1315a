"""Batched propagation with differentiation at the size of the workload:
64 uneven Henon-Heiles instances (D ~ U{4..12}, L ~ U{4..8}, numpy seed 7:
18.6 M packed rows, 392 M grid points) in one plan (one launch per H.psi),
dipole mu = sum_i q_i (evaluated on the device), a control field of 200
seeded piecewise-constant N(0, 0.01) amplitudes, Chebyshev steps of dt = 0.1
on the device (sg4_cheby_*), split-precision contractions.  The gradient of
J = sum_i |<target_i|psi_i(T)>|^2 with respect to the field amplitudes
(adjoint with checkpointed recomputation, sg4_cheby_field_gradient) is
wrapped in a torch.autograd.Function here, in the test harness only, and
checked against central finite differences (FP64, 1e-6 relative); the
split-precision gradient is compared with the FP64 one.  The horizon is cut
to the first segments of the field (5 steps each) to keep the test short:
the 1000-step run is timed by bench.py --propagation."""
import time

import numpy as np
import pytest

import elvibrot_amd as E
from elvibrot_amd.propagation import (ChebyshevPropagator, spectral_bounds, propagate_device,
                                      field_gradient_device)

pytestmark = pytest.mark.gpu

NSEG_TEST, STEPS = 2, 5


def config5(cuda):
    rng = np.random.default_rng(7)
    D = rng.integers(4, 13, size=64)
    L = rng.integers(4, 9, size=64)
    bases = [E.SG4Basis(int(d), int(l)) for d, l in zip(D, L)]
    op = E.SG4Batch(bases, max_nvec=4)
    eps = np.random.default_rng(7).normal(0.0, 0.01, 200)
    return bases, op, eps


def test_config5_field_gradient(cuda):
    import torch
    t0 = time.perf_counter()
    bases, op, eps_all = config5(cuda)
    nb = op.basis.nb_ba
    mu = op.grid_dipole()
    lo, hi = spectral_bounds(op)
    w = 0.05 * (hi - lo) + np.abs(eps_all).max() * mu.abs().max().item()
    print(f"64 instances: {nb} rows, {op.basis.sum_nq} grid points; set-up {time.perf_counter() - t0:.1f} s; "
          f"spectral range [{lo:.3f}, {hi:.3f}]", flush=True)
    blocks = list(zip(op.row0[:-1], op.row0[1:]))
    psi0 = torch.zeros(1, nb, dtype=torch.complex128, device=cuda)
    target = torch.zeros_like(psi0)
    for r0, r1 in blocks:          # per instance: start in the ground row, target (|0> + |1>)/sqrt 2
        psi0[0, r0] = 1.0
        target[0, r0] = np.sqrt(0.5)
        target[0, r0 + 1] = np.sqrt(0.5)
    eps0 = eps_all[:NSEG_TEST]

    props = {p: ChebyshevPropagator(op, 0.1, lo - w, hi + w, mu=mu, prec=p)
             for p in (E.PREC_FP64, E.PREC_F32X2)}

    class FieldObjective(torch.autograd.Function):
        """J(eps) with the device adjoint as its backward (test harness only)."""

        @staticmethod
        def forward(ctx, eps, prec):
            J, g = field_gradient_device(props[prec], psi0, target, eps.detach().cpu().numpy(),
                                         STEPS, checkpoint_every=1, blocks=blocks)
            ctx.save_for_backward(torch.tensor(g, dtype=torch.float64))
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            (g,) = ctx.saved_tensors
            return gJ * g, None

    grads = {}
    for p in (E.PREC_FP64, E.PREC_F32X2):
        e = torch.tensor(eps0, dtype=torch.float64, requires_grad=True)
        t1 = time.perf_counter()
        J = FieldObjective.apply(e, p)
        J.backward()
        grads[p] = (J.item(), e.grad.numpy().copy())
        print(f"prec {p}: J {J.item():.12f}, dJ/deps {e.grad.numpy()}, "
              f"{time.perf_counter() - t1:.1f} s for {NSEG_TEST * STEPS} steps "
              f"({props[p].nmax} orders per step) forward + adjoint", flush=True)

    def Jof(ev):
        pT = propagate_device(props[E.PREC_FP64], psi0, ev, STEPS)
        return sum(abs(torch.sum(target[0, r0:r1].conj() * pT[0, r0:r1]).item()) ** 2
                   for r0, r1 in blocks)

    # two FP64 runs with atomic scatter-adds differ by their order, amplified
    # by the Smolyak cancellation (1-4e-11 relative in J): divided by the
    # step h of the differences that is up to ~1e-6 of the gradient.  The
    # differences are therefore taken with the order-independent sum, whose
    # runs are bitwise reproducible, so they carry only the O(h^2) error.
    J64, g64 = grads[E.PREC_FP64]
    op.set_deterministic(True)
    assert abs(Jof(eps0) - J64) < 1e-9 * J64
    h = 1e-4
    for s in range(NSEG_TEST):
        ep, em = eps0.copy(), eps0.copy()
        ep[s] += h
        em[s] -= h
        fd = (Jof(ep) - Jof(em)) / (2 * h)
        print(f"segment {s}: adjoint {g64[s]:.10e} fd {fd:.10e}", flush=True)
        assert abs(g64[s] - fd) < 1e-6 * max(abs(fd), 1e-6)
    Jx, gx = grads[E.PREC_F32X2]
    rel = np.max(np.abs(gx - g64)) / np.max(np.abs(g64))
    print(f"split-precision vs FP64: gradient {rel:.2e}, J {abs(Jx - J64) / J64:.2e}; "
          f"total {time.perf_counter() - t0:.1f} s")
    # the split-precision contractions carry FP32-level errors, amplified by
    # the Smolyak cancellation on these smooth packets (started in the ground
    # rows: up to 1.2e5 at 10D L=8), and FP32 rounding of the FP64 states
    # makes them differ from run to run: J measured 1.0-1.4e-5 from FP64,
    # the gradient 1.2e-4
    assert rel < 1e-3 and abs(Jx - J64) < 1e-4 * J64
