"""The reference's two-surface propagations of pyrazine
(Working_tests/MPI_tests/12D_propagation_S3 at L = 1 and 24D_propagation_S3
at L = 2, the levels their run_jobs use) through SG4Surfaces and the
Chebyshev propagator: the packet starts in the
ground HO function of diabatic state 2 (its New_Read_WP0 line), and the
autocorrelation C(t) = <psi0|psi(t)> after every 0.1 fs (the reference's
micro times: WPdeltaT = 2 fs, nb_micro = 20) is compared with the 61 values
the reference stores (tests/golden/pyrazine{12d_L1,24d_L2}_autocorr_benchmark.txt:
T in au, Re C, Im C, |C|).  Times in au from the CODATA 2006 constants, the
reference's default set; its file's time grid is checked first."""
import os

import numpy as np
import pytest

from elvibrot_amd import psi_io
from elvibrot_amd.potentials import codata2006_au, pyrazine_basis, pyrazine_surfaces
from elvibrot_amd.propagation import ChebyshevPropagator, spectral_bounds

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("D,L", [(12, 1), (24, 2)])
def test_pyrazine_autocorrelation(cuda, golden, D, L):
    import torch
    ref = np.loadtxt(os.path.join(golden, f"pyrazine{D}d_L{L}_autocorr_benchmark.txt"))
    _, Ta = codata2006_au()
    dt = 0.1e-15 / Ta
    assert np.max(np.abs(ref[:, 0] - dt * np.arange(len(ref)))) < 1e-12 * ref[-1, 0]
    b = pyrazine_basis(D, L)
    S = pyrazine_surfaces(b)
    nb = b.nb_ba
    r0 = int(np.nonzero(~psi_io.packed_indices(b).any(axis=1))[0][0])
    print(f"{D}D L={L}: {nb} packed functions per state, {b.nterms} terms", flush=True)
    lo, hi = spectral_bounds(S)
    prop = ChebyshevPropagator(S, dt, lo, hi)
    x = torch.zeros(1, 2 * nb, dtype=torch.complex128, device=cuda)
    x[0, nb + r0] = 1.0
    C = [complex(x[0, nb + r0].item())]
    for _ in range(len(ref) - 1):
        x, _ = prop.step(x)
        C.append(complex(x[0, nb + r0].item()))
    C = np.array(C)
    err = np.abs(C - (ref[:, 1] + 1j * ref[:, 2]))
    print(f"autocorrelation over {len(ref)} times: max |C - C_ref| {err.max():.2e}, "
          f"|C(6 fs)| {abs(C[-1]):.6f} vs {ref[-1, 3]:.6f}")
    # the reference cuts its Chebyshev series at poly_tol = 1e-10 (|c_k| ||w_k||,
    # its shell_run); measured here: 1e-14 (12D), 9e-13 (24D)
    assert err.max() < 1e-11
    assert abs(np.linalg.norm(x[0].cpu().numpy()) - 1.0) < 1e-10
