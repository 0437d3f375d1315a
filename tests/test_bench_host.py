"""bench.py's host side on the CPU: the one-GPU guard, and the reference
algorithm's flop count that the bench line divides by the kernel time,
against a hand count and against the figures stated for the workloads from
the Smolyak rules (n(l) = 1 + 2l, L-D+1 <= |l| <= L), to their three
digits: F1 = 9.70 M per vector at 6D L=6, 4.80 G per vector and 76.8 G per
16-vector block at 10D L=8."""
import os
import subprocess
import sys

import pytest

import elvibrot_amd as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("D,L,nvec,expect", [(6, 6, 1, 9.70e6), (10, 8, 1, 4.80e9),
                                             (10, 8, 16, 76.8e9)])
def test_reference_flops_match_the_stated_counts(D, L, nvec, expect):
    assert bench.reference_flops(E.SG4Basis(D, L), nvec) == pytest.approx(expect, rel=1e-3)


def test_reference_flops_brute_force_small_set():
    """2D L=1: terms (0,0) w=-1, (1,0), (0,1) w=1 with n = 1 or 3 points:
    S_t = 2 N_t sum_k n_k -> (1,1): 2*1*2 = 4, (3,1) and (1,3): 2*3*4 = 24
    each; F1 = 3 (4 + 24 + 24) = 156 per vector."""
    assert bench.reference_flops(E.SG4Basis(2, 1), 1) == 156.0
