"""Host logic of the propagators (CPU only)."""
import math

import pytest

from elvibrot_amd.propagation import taylor_npoly


@pytest.mark.parametrize("alpha,tol,n", [(1.0, 1e-20, 22), (2.5, 1e-20, 30), (10.0, 1e-12, 47),
                                         (1.0, 1e-12, 15)])
def test_taylor_order_is_first_term_below_tol(alpha, tol, n):
    """cof_nOD's second loop ends at the first order whose term alpha^i/i! is
    below poly_tol: alpha^n/n! <= tol < alpha^(n-1)/(n-1)!."""
    assert taylor_npoly(alpha, tol) == n
    term = lambda i: alpha ** i / math.factorial(i)
    assert term(n) <= tol < term(n - 1)


def test_taylor_order_first_loop_runs_to_max_poly():
    """cof_nOD's first loop asks |log(sum_k<=i alpha^k/k!) - alpha| <= poly_tol; in
    FP64 the partial sum of e^0.1 never gets its logarithm within 1e-20 of 0.1,
    so the order is max_poly + 1 and only the per-step norm cut stops the
    series, as in the reference."""
    assert taylor_npoly(0.1, 1e-20, max_poly=5000) == 5001
    assert taylor_npoly(0.1, 1e-20, max_poly=50) == 51
