"""Split-precision MFMA mode (SG4_PREC_F32X2) on the MI355X through the C ABI.

The contractions run on the matrix cores with three bf16 limbs per FP32
operand (six products, FP32 accumulation), each as accurate as an FP32
contraction; the kinetic part, the sum over terms and the result are FP64.
The Smolyak combination amplifies the per-term rounding: the potential
parts W_t r_t of the terms sum to H.psi with cancellations of up to
max_e sum_t |W_t r_t| / max|H psi| ~ 2e2 (6D L=6) and 7e2 (10D L=5).  Stated
tolerance of the mode against the FP64 oracle, checked below:
max|got - ref| <= 8 eps32 max_e sum_t |W_t r_t| (the oracle computes the
bound), and a 2-norm relative error <= 2e-6.
"""
import numpy as np
import pytest

import elvibrot_amd as E
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL2 = 2e-6                               # 2-norm relative
EPS32 = float(np.finfo(np.float32).eps)
KCOND = 8.0                               # max-norm: KCOND eps32 max_e sum_t |W_t r_t|


def errs(got, ref):
    return (np.linalg.norm(got - ref) / np.linalg.norm(ref),
            np.max(np.abs(got - ref)) / np.max(np.abs(ref)))


def cond_ratio(got, ref, o, psi, V, D):
    """max|got - ref| in units of eps32 max_e sum_t |W_t r_t| (potential parts)."""
    bound = o.absbound(psi, V, np.zeros(D))
    return np.max(np.abs(got - ref)) / (EPS32 * np.max(bound))


@pytest.mark.parametrize("D,L,nvec", [(6, 3, 4), (6, 6, 16), (10, 5, 16), (4, 5, 7), (3, 10, 5),
                                      (2, 10, 16), (21, 2, 3)])
def test_f32x2_parity_vs_oracle(cuda, D, L, nvec):
    print(f"case D={D} L={L} nvec={nvec}", flush=True)
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=16, potential=("henon_heiles", 0.111803))
    assert op.split_info()[0]
    o = O.SG4(D, L)
    V = o.grid_hh(0.111803)
    psi = np.random.default_rng(100 * D + L).standard_normal((b.nb_ba, nvec))
    ref = o.hpsi(psi, V, -0.5 * np.ones(D))
    got = op.apply_host(psi, E.PREC_F32X2)
    e2, em = errs(got, ref)
    rc = cond_ratio(got, ref, o, psi, V, D)
    f, ngm = op.mfma_stats()
    print(f"D={D} L={L}: 2-norm {e2:.2e} max {em:.2e}, max / (eps32 bound) {rc:.3f}; "
          f"group matrices {ngm}, MFMA flops {f:.3g}")
    assert e2 < TOL2 and rc < KCOND
    assert ngm > 0 and f > 0


def test_f32x2_general_potential_absent_rows(cuda):
    # arbitrary grid values, unequal kinetic coefficients, LB < LG (absent map rows)
    D, L, LB = 4, 5, 3
    b = E.SG4Basis(D, L, LB)
    o = O.SG4(D, L, LB)
    rng = np.random.default_rng(11)
    V = rng.standard_normal(o.sum_nq)
    ck = np.array([-0.5, -0.25, -1.0, -0.75])
    op = E.SG4Operator(b, 8, potential=V, ck=ck)
    if not op.split_info()[0]:
        pytest.skip("this basis does not take the split path")
    psi = rng.standard_normal((b.nb_ba, 6))
    got, ref = op.apply_host(psi, E.PREC_F32X2), o.hpsi(psi, V, ck)
    e2, em = errs(got, ref)
    rc = cond_ratio(got, ref, o, psi, V, D)
    print(f"4D L=5 LB=3 random V: 2-norm {e2:.2e} max {em:.2e} ratio {rc:.3f}")
    assert e2 < TOL2 and rc < KCOND


@pytest.mark.parametrize("D,L,nvec", [(6, 6, 32), (6, 6, 24), (4, 5, 32), (10, 4, 32), (4, 5, 64), (6, 4, 48)])
def test_f32x2_wide_blocks(cuda, D, L, nvec):
    """Blocks of 24 to 64 vectors (work items of up to 64 vectors where the
    LDS budget allows) at the stated tolerance of the mode."""
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=64, potential=("henon_heiles", 0.111803))
    o = O.SG4(D, L)
    V = o.grid_hh(0.111803)
    psi = np.random.default_rng(7 * D + L + nvec).standard_normal((b.nb_ba, nvec))
    got, ref = op.apply_host(psi, E.PREC_F32X2), o.hpsi(psi, V, -0.5 * np.ones(D))
    e2, em = errs(got, ref)
    rc = cond_ratio(got, ref, o, psi, V, D)
    print(f"D={D} L={L} nvec={nvec}: 2-norm {e2:.2e} max {em:.2e} ratio {rc:.3f}")
    assert e2 < TOL2 and rc < KCOND


def test_f32x2_batch_of_uneven_instances(cuda):
    """One batched plan over instances of different D and L (the layout of the
    64-instance propagation batch): each instance's rows of the split-precision
    H.psi against the oracle of that instance, at the tolerance above."""
    specs = [(4, 5), (6, 6), (5, 4), (8, 5), (3, 8)]
    bases = [E.SG4Basis(D, L) for D, L in specs]
    op = E.SG4Batch(bases, max_nvec=4)
    assert op.split_info()[0]
    psi = np.random.default_rng(12).standard_normal((op.basis.nb_ba, 4))
    got = op.apply_host(psi, E.PREC_F32X2)
    for (D, L), r0, r1 in zip(specs, op.row0[:-1], op.row0[1:]):
        o = O.SG4(D, L)
        V = o.grid_hh(0.111803)
        ref = o.hpsi(psi[r0:r1], V, -0.5 * np.ones(D))
        e2, em = errs(got[r0:r1], ref)
        rc = cond_ratio(got[r0:r1], ref, o, psi[r0:r1], V, D)
        print(f"instance D={D} L={L}: 2-norm {e2:.2e} max {em:.2e} ratio {rc:.3f}")
        assert e2 < TOL2 and rc < KCOND


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_F32X2])
def test_narrow_blocks_take_the_launch_set_for_16(cuda, prec):
    """With only the 16-vector launch set (the pinned launches of the 10D
    workload), blocks of 1..16 vectors run with it (sg4_apply pads them to 2,
    4, 8 or 16 vectors; sg4_plan_config reports the launch) and keep the
    parity of the mode: 1e-12 in FP64, the split-precision bound otherwise."""
    D, L = 6, 6
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=16, potential=("henon_heiles", 0.111803))
    pinned = {E.PREC_FP64: (512, 163840, 2048, 3), E.PREC_F32X2: (1024, 163840, 512, 2)}[prec]
    op.set_config(16, prec, *pinned)
    o = O.SG4(D, L)
    V = o.grid_hh(0.111803)
    rng = np.random.default_rng(21)
    for nvec in (1, 3, 5, 9, 16):
        c = op.config(nvec, prec)
        assert (c["nthreads"], c["lds_budget"], c["grid"], c["xcd"]) == pinned
        psi = rng.standard_normal((b.nb_ba, nvec))
        got, ref = op.apply_host(psi, prec), o.hpsi(psi, V, -0.5 * np.ones(D))
        e2, em = errs(got, ref)
        print(f"prec {prec} nvec {nvec}: 2-norm {e2:.2e} max {em:.2e}")
        if prec == E.PREC_FP64:
            assert em < 1e-12
        else:
            assert e2 < TOL2 and cond_ratio(got, ref, o, psi, V, D) < KCOND


def test_f32x2_10d_L8_block16_against_fp64(cuda):
    """The headline workload: every vector of the 16-block against the FP64
    apply (itself at 1e-12 of the oracle) in the 2-norm, two of them against
    the oracle with the conditioning bound of the max norm, symmetry of the
    projected matrix, and linearity."""
    import torch
    D, L, nvec = 10, 8, 16
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=nvec)
    g = torch.Generator(device=cuda).manual_seed(2026)
    X = torch.randn(b.nb_ba, nvec, device=cuda, dtype=torch.float64, generator=g)
    X, _ = torch.linalg.qr(X)
    Xc = X.T.contiguous()
    H64 = torch.empty_like(Xc)
    H32 = torch.empty_like(Xc)
    op.apply(Xc, H64, E.PREC_FP64)
    op.apply(Xc, H32, E.PREC_F32X2)
    torch.cuda.synchronize()
    e2 = (torch.linalg.norm(H32 - H64, dim=1) / torch.linalg.norm(H64, dim=1)).max().item()
    em = (torch.max(torch.abs(H32 - H64)) / torch.max(torch.abs(H64))).item()
    S = Xc @ H32.T
    sym = (torch.max(torch.abs(S - S.T)) / torch.max(torch.abs(S))).item()
    Y = (2 * Xc[0:1] - 3 * Xc[1:2]).contiguous()
    HY = torch.empty_like(Y)
    op.apply(Y, HY, E.PREC_F32X2)
    lin = 2 * H32[0:1] - 3 * H32[1:2]
    el = (torch.linalg.norm(HY - lin) / torch.linalg.norm(lin)).item()
    o = O.SG4(D, L)
    V = o.grid_hh(0.111803)
    psi = Xc[:2].T.cpu().numpy()
    got = H32[:2].T.cpu().numpy()
    ref = o.hpsi(psi, V, -0.5 * np.ones(D))
    rc = cond_ratio(got, ref, o, psi, V, D)
    print(f"10D L=8 x16 split vs FP64: worst vector 2-norm {e2:.2e}, max {em:.2e}, "
          f"symmetry {sym:.2e}, linearity {el:.2e}; 2 vectors vs oracle: max / (eps32 bound) "
          f"{rc:.3f}")
    assert e2 < TOL2 and rc < KCOND
    assert sym < TOL2 and el < TOL2


def test_f32x2_repeatable_and_graph(cuda):
    """Two applies differ only by the order of the FP64 atomic additions; a
    captured graph of the split mode equals the eager apply to that level."""
    import torch
    D, L, nvec = 6, 5, 8
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=nvec)
    X = torch.randn(b.nb_ba, nvec, device=cuda, dtype=torch.float64,
                    generator=torch.Generator(device=cuda).manual_seed(5))
    Y1, Y2, Yg = (torch.empty_like(X) for _ in range(3))
    op.apply_rows(X, Y1, E.PREC_F32X2)
    op.apply_rows(X, Y2, E.PREC_F32X2)
    for _ in range(3):
        op.apply_graph(X, Yg, E.PREC_F32X2)
        torch.cuda.synchronize()
        assert (torch.max(torch.abs(Yg - Y1)) / torch.max(torch.abs(Y1))).item() < 1e-12
    assert (torch.max(torch.abs(Y2 - Y1)) / torch.max(torch.abs(Y1))).item() < 1e-12


def test_f32x2_rejected_off_the_split_path(cuda):
    # one pass per mode (grouping off) is not the split path: the mode is refused
    b = E.SG4Basis(4, 4)
    op = E.SG4Operator(b, max_nvec=2)
    op.set_grouping(False)
    with pytest.raises(E.SG4Error):
        op.apply_host(np.ones((b.nb_ba, 2)), E.PREC_F32X2)


def test_f32x2_deterministic_bitwise(cuda):
    """With the deterministic sum the split-precision mode is bitwise
    repeatable, also across launch configurations (the per-term MFMA results
    do not depend on the item order or the LDS budget)."""
    import torch
    D, L, nvec = 6, 6, 16
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=nvec)
    op.set_deterministic(True)
    X = torch.randn(b.nb_ba, nvec, device=cuda, dtype=torch.float64,
                    generator=torch.Generator(device=cuda).manual_seed(21))
    Y1, Y2, Y3 = (torch.empty_like(X) for _ in range(3))
    op.apply_rows(X, Y1, E.PREC_F32X2)
    op.apply_rows(X, Y2, E.PREC_F32X2)
    op.set_config(nvec, E.PREC_F32X2, nthreads=512, lds_budget=163840, grid=1024, xcd=3)
    op.apply_rows(X, Y3, E.PREC_F32X2)
    torch.cuda.synchronize()
    assert torch.equal(Y1, Y2) and torch.equal(Y1, Y3)
