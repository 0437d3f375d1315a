"""FP64 H.psi on the MI355X through the C ABI, against the oracle (parity
bar: 1e-12 relative to max|H psi|), plus size-independent properties at the
10D L=8 block-16 size."""
import numpy as np
import pytest

import elvibrot_amd as E
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def hh_pair(D, L, max_nvec=16):
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=max_nvec, potential=("henon_heiles", 0.111803))
    o = O.SG4(D, L)
    V = o.grid_hh(0.111803)
    return b, op, o, V


@pytest.mark.parametrize("D,L,nvec", [(6, 3, 1), (6, 3, 5), (6, 6, 16), (21, 2, 3),
                                      (4, 5, 7), (3, 10, 2), (2, 12, 13)])
def test_fp64_parity_host_entry(cuda, D, L, nvec):
    print(f"case D={D} L={L} nvec={nvec}", flush=True)
    b, op, o, V = hh_pair(D, L)
    rng = np.random.default_rng(12345 + D * 100 + L)
    psi = rng.standard_normal((b.nb_ba, nvec))
    ref = o.hpsi(psi, V, -0.5 * np.ones(D))
    got = op.apply_host(psi)
    assert rel(got, ref) < TOL


@pytest.mark.parametrize("D,L", [(1, 0), (1, 3), (1, 6), (1, 15), (3, 0), (5, 0), (2, 1), (2, 12)])
@pytest.mark.parametrize("nvec", [3, 17])
def test_smallest_sets_all_precisions(cuda, D, L, nvec):
    """The edge of the term sets: one mode (the harmonic oscillator,
    diag(n + 1/2); 31 points at L = 15, the widest single-mode tile of the
    matrix cores), L = 0 (one function on the one-point grid: D/2), 2D L = 1,
    and 2D L = 12 (modes of 25 points).  Modes of more than 24 points take
    the grid path in FP64 and FP32 and the split algebra on the matrix cores.
    FP64 to the parity bar and the closed form; the reduced modes within 2e-6
    (2-norm) of the oracle."""
    print(f"case D={D} L={L} nvec={nvec}", flush=True)
    b, op, o, V = hh_pair(D, L, max_nvec=32)
    split, nnz = op.split_info()
    assert nnz > 0 and split == (2 * L + 1 <= 24)
    rng = np.random.default_rng(7 + 10 * D + L)
    psi = rng.standard_normal((b.nb_ba, nvec))
    ref = o.hpsi(psi, V, -0.5 * np.ones(D))
    got = op.apply_host(psi)
    assert rel(got, ref) < TOL
    if o.nterms == 1:
        diag = np.full(b.nb_ba, D / 2) if L == 0 else np.arange(b.nb_ba) + 0.5
        assert rel(got, diag[:, None] * psi) < TOL
    for prec in (E.PREC_F32, E.PREC_F32X2):
        g = op.apply_host(psi, prec)
        e2 = np.linalg.norm(g - ref) / np.linalg.norm(ref)
        print(f"prec {prec}: 2-norm {e2:.2e}")
        assert e2 < 2e-6


def test_fp64_parity_device_entries(cuda):
    import torch
    D, L, nvec = 6, 6, 16
    b, op, o, V = hh_pair(D, L)
    psi = np.random.default_rng(2026).standard_normal((b.nb_ba, nvec))
    ref = o.hpsi(psi, V, -0.5 * np.ones(D))
    x = torch.tensor(psi.T.copy(), device=cuda)           # (nvec, nb_ba) = column-major block
    y = torch.empty_like(x)
    op.apply(x, y)
    torch.cuda.synchronize()
    assert rel(y.cpu().numpy().T, ref) < TOL
    xr = torch.tensor(psi, device=cuda)                   # rows layout (nb_ba, nvec)
    yr = torch.empty_like(xr)
    op.apply_rows(xr, yr)
    torch.cuda.synchronize()
    assert rel(yr.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("max_nvec,nvec", [(12, 9), (12, 12), (32, 20), (16, 5), (4, 3), (40, 33)])
def test_padded_block_widths(cuda, max_nvec, nvec):
    """sg4_apply stages a block padded with zero vectors to the next power of
    two (<= 16) or multiple of 16 when max_nvec allows (12 -> 16 does not:
    unpadded), and returns exactly the nvec columns asked for."""
    import torch
    D, L = 5, 4
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=max_nvec)
    o = O.SG4(D, L)
    psi = np.random.default_rng(max_nvec * 100 + nvec).standard_normal((b.nb_ba, nvec))
    ref = o.hpsi(psi, o.grid_hh(0.111803), -0.5 * np.ones(D))
    x = torch.tensor(psi.T.copy(), device=cuda)
    ybuf = torch.full((nvec + 1, b.nb_ba), 7.0, device=cuda, dtype=torch.float64)
    op.apply(x, ybuf[:nvec])
    torch.cuda.synchronize()
    assert rel(ybuf[:nvec].cpu().numpy().T, ref) < TOL
    assert torch.all(ybuf[nvec] == 7.0)          # nothing written past the block


def test_fp64_host_potential_equals_device_potential(cuda):
    D, L = 5, 4
    b = E.SG4Basis(D, L)
    o = O.SG4(D, L)
    V = o.grid_hh(0.111803)
    op_dev = E.SG4Operator(b, 4, potential=("henon_heiles", 0.111803))
    op_host = E.SG4Operator(b, 4, potential=V, ck=-0.5 * np.ones(D))
    psi = np.random.default_rng(3).standard_normal((b.nb_ba, 4))
    assert rel(op_dev.apply_host(psi), op_host.apply_host(psi)) < 1e-14


def test_fp64_general_potential_and_coefficients(cuda):
    # arbitrary grid values and unequal kinetic coefficients, LB < LG (absent map rows)
    D, L, LB = 4, 5, 3
    b = E.SG4Basis(D, L, LB)
    o = O.SG4(D, L, LB)
    rng = np.random.default_rng(11)
    V = rng.standard_normal(o.sum_nq)
    ck = np.array([-0.5, -0.25, -1.0, -0.75])
    op = E.SG4Operator(b, 8, potential=V, ck=ck)
    psi = rng.standard_normal((b.nb_ba, 6))
    assert rel(op.apply_host(psi), o.hpsi(psi, V, ck)) < TOL


@pytest.mark.parametrize("deterministic", [False, True])
def test_10d_L8_block16_parity_extended_oracle(cuda, deterministic):
    """All 16 vectors of the 10D L=8 workload against the oracle evaluated in
    extended precision (long double contractions and sums, so the comparison
    measures the GPU's error alone): 2-norm relative error <= 1e-12 for every
    vector, and the max norm within 4 eps of max_e sum_t |W_t r_t| over the
    potential parts of the terms (their FP64 contraction rounding, amplified
    by the Smolyak combination; the kinetic part is exact on the split path).
    With the deterministic sum two applies are bitwise identical and the max
    norm is within 1e-12 too (the order-dependent FP64 atomic rounding is
    gone: measured 1.8e-15 in the 2-norm, 4.0e-14 in the max norm, against
    5.2e-14 and 6.1e-12 with FP64 atomics)."""
    D, L, nvec = 10, 8, 16
    b, op, o, V = hh_pair(D, L, nvec)
    op.set_deterministic(deterministic)
    psi = np.random.default_rng(2026).standard_normal((b.nb_ba, nvec))
    got = op.apply_host(psi)
    if deterministic:
        assert np.array_equal(got, op.apply_host(psi))
    ref = o.hpsi_ld(psi, V, -0.5 * np.ones(D))
    bound = o.absbound(psi, V, np.zeros(D))
    eps = np.finfo(np.float64).eps
    e2 = np.max(np.linalg.norm(got - ref, axis=0) / np.linalg.norm(ref, axis=0))
    r = np.max(np.abs(got - ref)) / (eps * np.max(bound))
    print(f"10D L=8 x16 deterministic={deterministic}: worst 2-norm {e2:.2e}, max-norm rel "
          f"{rel(got, ref):.2e}, max|err| / (eps max bound) {r:.2f} "
          f"(bound / max|H psi| = {np.max(bound) / np.max(np.abs(ref)):.3g})")
    assert e2 < TOL
    assert r < 4
    if deterministic:    # exact sum of the contributions: measured 4e-14 in the max norm
        assert rel(got, ref) < TOL


def test_10d_L8_block16_properties(cuda):
    """The 10D L=8 block-16 workload: symmetry <phi|H psi> = <H phi|psi>, linearity, and
    agreement of every vector's result with the one-vector apply."""
    import torch
    D, L, nvec = 10, 8, 16
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=nvec)
    g = torch.Generator(device=cuda).manual_seed(7)
    X = torch.randn(nvec, b.nb_ba, device=cuda, dtype=torch.float64, generator=g)
    HX = torch.empty_like(X)
    op.apply(X, HX)
    S = X @ HX.T
    assert (torch.max(torch.abs(S - S.T)) / torch.max(torch.abs(S))).item() < 1e-12
    # linearity: H(2 x0 - 3 x1) = 2 H x0 - 3 H x1 (2-norm: the max norm is
    # limited by the conditioning of the Smolyak sum, see the parity test)
    nrm = lambda a: torch.linalg.norm(a).item()
    Y = (2 * X[0:1] - 3 * X[1:2]).contiguous()
    HY = torch.empty_like(Y)
    op.apply(Y, HY)
    lin = 2 * HX[0:1] - 3 * HX[1:2]
    assert nrm(HY - lin) / nrm(lin) < 1e-12
    # a single vector through the one-vector path equals its column of the block
    H3 = torch.empty_like(X[3:4])
    op.apply(X[3:4].contiguous(), H3)
    assert nrm(H3 - HX[3:4]) / nrm(HX[3:4]) < 1e-12


def test_12d_L8_largest_range_instance_properties(cuda):
    """The largest instance of the D x L range (12D, L=8: 125 970 terms,
    120.2 M grid points, nb_ba 4 673 345), where the oracle would take hours:
    the split path is taken, H is symmetric and linear, a vector applied alone
    equals its column of the block, and the FP32 mode agrees with FP64 to its
    precision."""
    import torch
    D, L, nvec = 12, 8, 4
    b = E.SG4Basis(D, L)
    assert (b.nterms, b.nb_ba) == (125970, 4673345)
    op = E.SG4Operator(b, max_nvec=nvec)
    assert op.split_info()[0]
    g = torch.Generator(device=cuda).manual_seed(12)
    X = torch.randn(nvec, b.nb_ba, device=cuda, dtype=torch.float64, generator=g)
    HX = torch.empty_like(X)
    op.apply(X, HX)
    nrm = lambda a: torch.linalg.norm(a).item()
    S = X @ HX.T
    assert (torch.max(torch.abs(S - S.T)) / torch.max(torch.abs(S))).item() < 1e-12
    Y = (0.5 * X[0:1] + 4 * X[2:3]).contiguous()
    HY = torch.empty_like(Y)
    op.apply(Y, HY)
    lin = 0.5 * HX[0:1] + 4 * HX[2:3]
    assert nrm(HY - lin) / nrm(lin) < 1e-12
    H1 = torch.empty_like(X[1:2])
    op.apply(X[1:2].contiguous(), H1)
    assert nrm(H1 - HX[1:2]) / nrm(HX[1:2]) < 1e-12
    H32 = torch.empty_like(X)
    op.apply(X, H32, E.PREC_F32)
    r32 = nrm(H32 - HX) / nrm(HX)
    print(f"12D L8: FP32 vs FP64 2-norm rel {r32:.2e}")
    assert r32 < 1e-5


def test_errors_are_reported(cuda):
    b, op, o, V = hh_pair(3, 2, max_nvec=2)
    with pytest.raises(E.SG4Error):
        op.apply_host(np.zeros((b.nb_ba, 3)))
    with pytest.raises(E.SG4Error):
        op.apply_host(np.zeros((b.nb_ba, 1)), prec=7)
    # psi and hpsi must be distinct blocks (the kernels read one while adding
    # into the other): in place, or shifted by one row, is rejected
    import torch
    x = torch.zeros(3, b.nb_ba, device=cuda, dtype=torch.float64)
    with pytest.raises(E.SG4Error, match="overlap"):
        op.apply(x[0:2], x[0:2])
    with pytest.raises(E.SG4Error, match="overlap"):
        op.apply_rows(x[0:2].reshape(b.nb_ba, 2), x[1:3].reshape(b.nb_ba, 2))
    y = torch.empty_like(x[0:2])
    op.apply(x[0:2], y)                   # adjacent, distinct blocks are fine
    op.apply(x[1:2], x[0:1])


def test_mode_beyond_the_matrix_core_tile(cuda):
    """2D L=16: modes of 33 points, more than a matrix-core tile holds.  The
    split-precision mode refuses the plan with the reason (before any
    launch); FP64 and FP32 run it on the grid path, FP64 to the parity bar."""
    b, op, o, V = hh_pair(2, 16, max_nvec=4)
    psi = np.random.default_rng(33).standard_normal((b.nb_ba, 3))
    ref = o.hpsi(psi, V, -0.5 * np.ones(2))
    assert rel(op.apply_host(psi), ref) < TOL
    g32 = op.apply_host(psi, E.PREC_F32)
    assert np.linalg.norm(g32 - ref) / np.linalg.norm(ref) < 2e-6
    with pytest.raises(E.SG4Error, match="more than 32 grid points"):
        op.apply_host(psi, E.PREC_F32X2)
    assert rel(op.apply_host(psi), ref) < TOL          # the plan is still usable


@pytest.mark.parametrize("D,L", [(6, 6), (10, 5)])
def test_f32_mode_accuracy(cuda, D, L):
    """FP32 mode contractions (FP64 gather/scatter) against the FP64 oracle:
    max|err| <= eps32 max_e sum_t |W_t r_t| (the conditioning scale of the
    Smolyak sum; measured 0.08-0.4 of it) and 2-norm relative error < 1e-5
    (measured 6e-7 - 9e-7)."""
    b, op, o, V = hh_pair(D, L)
    psi = np.random.default_rng(99).standard_normal((b.nb_ba, 4))
    ref = o.hpsi(psi, V, -0.5 * np.ones(D))
    bound = o.absbound(psi, V, -0.5 * np.ones(D))
    got = op.apply_host(psi, prec=E.PREC_F32)
    err2 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    eps32 = float(np.finfo(np.float32).eps)
    r = np.max(np.abs(got - ref)) / (eps32 * np.max(bound))
    print(f"f32 max-norm rel {rel(got, ref):.2e}, 2-norm rel {err2:.2e}, "
          f"max|err| / (eps32 max bound) {r:.3f}")
    assert r < 1
    assert err2 < 1e-5
    assert rel(got, ref) > 1e-12   # it really ran in reduced precision


@pytest.mark.parametrize("L,nvec", [(4, 3), (10, 5)])
def test_type10_jacobi_parity(cuda, L, nvec):
    import elvibrot_amd.potentials as P
    op = P.jacobi_triatomic_operator(L, max_nvec=8)
    b = op.basis
    sc = np.sqrt(P.AMU * np.sqrt(0.4 / P.AMU))
    o = O.SG4(3, L, kind=[0, 0, 1], Q0=[3.0, 2.0, 0.0], scaleQ=[sc, sc, 1.0], nq_A=2, nb_A=2)
    V, G, Jac, sqRJ = P.jacobi_triatomic_grids(b)
    psi = np.random.default_rng(L).standard_normal((b.nb_ba, nvec))
    ref = o.hpsi(psi, V, None, type_op=10, G=G, Jac=Jac, sqRJ=sqRJ)
    got = op.apply_host(psi)
    assert rel(got, ref) < TOL


@pytest.mark.parametrize("D,L,nvec", [(2, 8, 3), (3, 5, 4), (5, 4, 2)])
def test_type10_nonsymmetric_G_general_grids_parity(cuda, D, L, nvec):
    """Type 10 with host grids in which nothing is special: G a random
    non-symmetric D x D matrix per point, Jac and sqRhoOVERJac random positive
    (the oracle's type-10 form is pinned for this case against its
    product-rule expansion)."""
    print(f"case D={D} L={L} nvec={nvec}", flush=True)
    b = E.SG4Basis(D, L)
    o = O.SG4(D, L)
    rng = np.random.default_rng(70 + D)
    n = o.sum_nq
    V = rng.standard_normal(n)
    G = rng.standard_normal(n * D * D)
    Jac = 0.5 + rng.random(n)
    sqRJ = 0.5 + rng.random(n)
    op = E.SG4Operator(b, 8, potential=V, type_op=10, G=G, Jac=Jac, sqRJ=sqRJ)
    psi = rng.standard_normal((b.nb_ba, nvec))
    ref = o.hpsi(psi, V, None, type_op=10, G=G, Jac=Jac, sqRJ=sqRJ)
    assert rel(op.apply_host(psi), ref) < TOL
    # the transposed G is a different operator, as the device computes it too
    Gt = G.reshape(n, D, D).transpose(0, 2, 1).ravel()
    op_t = E.SG4Operator(b, 8, potential=V, type_op=10, G=Gt, Jac=Jac, sqRJ=sqRJ)
    assert rel(op_t.apply_host(psi), ref) > 1e-3


def test_autotuned_config_keeps_parity(cuda):
    import torch
    D, L, nvec = 6, 6, 16
    b, op, o, V = hh_pair(D, L)
    ms = op.autotune(nvec)
    cfg = op.config(nvec)
    print("autotune", ms, cfg)
    assert ms > 0 and cfg["nthreads"] in (256, 512)
    psi = np.random.default_rng(4).standard_normal((b.nb_ba, nvec))
    ref = o.hpsi(psi, V, -0.5 * np.ones(D))
    assert rel(op.apply_host(psi), ref) < TOL


def test_graph_replays_match_eager(cuda):
    """Captured graph (fill kernel + H.psi kernel) replayed several times with
    eager work in between; every replay equals the eager result."""
    import torch
    D, L, nvec = 6, 5, 8
    b, op, o, V = hh_pair(D, L)
    g = torch.Generator(device=cuda).manual_seed(3)
    X = torch.randn(b.nb_ba, nvec, device=cuda, dtype=torch.float64, generator=g)
    Y = torch.empty_like(X)
    ref = torch.empty_like(X)
    op.apply_rows(X, ref)
    for r in range(4):
        Y.fill_(float("nan"))
        op.apply_graph(X, Y)
        torch.cuda.synchronize()
        assert torch.max(torch.abs(Y - ref)).item() <= 1e-13 * torch.max(torch.abs(ref)).item()
        Z = torch.empty_like(X)
        op.apply_rows(X * (r + 1), Z)          # eager work between replays
        torch.cuda.synchronize()
        assert torch.max(torch.abs(Z - (r + 1) * ref)).item() <= 1e-12 * torch.max(torch.abs(Z)).item()


@pytest.mark.parametrize("D,L,nvec,prec", [(6, 4, 5, E.PREC_FP64), (6, 4, 5, E.PREC_F32X2),
                                           (10, 8, 16, E.PREC_FP64), (10, 8, 16, E.PREC_F32X2)])
def test_roofline_numerators_from_the_term_set(cuda, D, L, nvec, prec):
    """The algorithmic flops and bytes per launch that the bench line divides
    by the kernel time (sg4_plan_stats after an apply), recounted from the
    basis tables: on the split path every term costs per vector 4 N_t n_k
    flops per active mode (n_k > 1: a basis->grid and a grid->basis product)
    plus N_t for the potential; the compulsory bytes are the FP64 potential
    (8 per grid point), the int32 map (4 per term row) and one read of psi
    plus one write of H.psi.  At 10D L=8 x16: 42.29 GFLOP and 0.699 GB."""
    import torch
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=nvec)
    if (D, L) == (10, 8):      # the launches the bench pins: no item beyond the LDS budget
        op.set_config(nvec, E.PREC_FP64, 512, 163840, 2048, 3)
        op.set_config(nvec, E.PREC_F32X2, 1024, 163840, 512, 2)
    x = torch.randn(nvec, b.nb_ba, device=cuda, dtype=torch.float64)
    y = torch.empty_like(x)
    op.apply(x, y, prec)
    torch.cuda.synchronize()
    st = op.stats(nvec)
    nq = {(k, l): b.basis1d(k, l)["nq"] for k in range(D) for l in range(L + 1)}
    n = np.array([[nq[k, int(l)] for k, l in enumerate(row)] for row in b.levels()], dtype=np.float64)
    N = np.prod(n, axis=1)
    flops = nvec * np.sum(4.0 * N * np.where(n > 1, n, 0.0).sum(axis=1) + N)
    nbytes = 8.0 * N.sum() + 4.0 * N.sum() + 16.0 * b.nb_ba * nvec
    print(f"D={D} L={L} x{nvec} prec {prec}: {st['flops']:.6g} flops, {st['bytes']:.6g} bytes")
    assert st["flops"] == pytest.approx(flops, rel=1e-12)
    assert st["bytes"] == pytest.approx(nbytes, rel=1e-12)
    if (D, L, nvec) == (10, 8, 16):
        assert flops == pytest.approx(42.29e9, rel=1e-3) and nbytes == pytest.approx(0.699e9, rel=1e-3)


def test_graph_replay_10d_L8_block16(cuda):
    """The headline block through one captured HIP graph (kinetic rows, term
    kernel and, with the deterministic sum, its clear and finish kernels):
    FP64 replays with eager work between them equal the eager apply bit for
    bit; the split-precision replays equal their eager apply to FP64 rounding
    of the atomic sum (max-norm 1e-10 relative)."""
    import torch
    D, L, nvec = 10, 8, 16
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, nvec, potential=("henon_heiles", 0.111803))
    g = torch.Generator(device=cuda).manual_seed(31)
    X = torch.randn(b.nb_ba, nvec, device=cuda, dtype=torch.float64, generator=g)
    op.set_deterministic(True)
    ref = torch.empty_like(X)
    op.apply_rows(X, ref)
    Y = torch.empty_like(X)
    for r in range(3):
        Y.fill_(float("nan"))
        op.apply_graph(X, Y)
        torch.cuda.synchronize()
        assert torch.equal(Y, ref)
        Z = torch.empty_like(X)
        op.apply_rows(2.0 * X, Z)                  # eager work between replays
        torch.cuda.synchronize()
        # (each contribution is truncated at 2^-64, which does not commute
        # with the doubling: equal to rounding, not bitwise)
        assert torch.max(torch.abs(Z - 2.0 * ref)).item() <= 1e-13 * torch.max(torch.abs(Z)).item()
    op.set_deterministic(False)
    ref32 = torch.empty_like(X)
    op.apply_rows(X, ref32, E.PREC_F32X2)
    for r in range(3):
        Y.fill_(float("nan"))
        op.apply_graph(X, Y, E.PREC_F32X2)
        torch.cuda.synchronize()
        err = torch.max(torch.abs(Y - ref32)).item() / torch.max(torch.abs(ref32)).item()
        assert err < 1e-10
    assert (torch.linalg.norm(ref32 - ref) / torch.linalg.norm(ref)).item() < 2e-6


@pytest.mark.parametrize("D,L,prec", [(6, 6, E.PREC_FP64), (4, 6, E.PREC_FP64), (6, 6, E.PREC_F32)])
def test_grouped_and_per_mode_passes_agree(cuda, D, L, prec):
    """Register-blocked mode groups (default) and one pass per mode give the
    same operator.  The two paths round differently and add the terms in a
    different order, and the Smolyak sum amplifies rounding by
    max_e sum_t |W_t r_t| / max|H psi|, so FP64 agreement is checked as
    max|g - u| <= 16 eps max_e sum_t |W_t r_t| (and 2-norm < 1e-12); FP32 to
    1e-5 in the 2-norm."""
    b, op, o, V = hh_pair(D, L)
    psi = np.random.default_rng(21).standard_normal((b.nb_ba, 5))
    g = op.apply_host(psi, prec)
    op.set_grouping(False)
    u = op.apply_host(psi, prec)
    op.set_grouping(True)
    if prec == E.PREC_FP64:
        bound = o.absbound(psi, V, -0.5 * np.ones(D))
        eps = np.finfo(np.float64).eps
        r = np.max(np.abs(g - u)) / (eps * np.max(bound))
        print(f"grouped vs per-mode: 2-norm {rel(g, u):.2e}, max|diff| / (eps max bound) {r:.2f}")
        assert r < 16 and rel(g, u) < 1e-12
        assert rel(g, o.hpsi(psi, V, -0.5 * np.ones(D))) < TOL
    else:
        assert rel(g, u) < 1e-5


@pytest.mark.parametrize("D,L,ck_scale", [(4, 5, 1.0), (6, 4, 0.7), (3, 9, 1.3)])
def test_split_path_in_use_and_matches_oracle(cuda, D, L, ck_scale):
    """HO bases with nq = nb on the complete Smolyak set take the split path
    (kinetic part as one sparse matrix); with mode-dependent kinetic
    coefficients and a random potential it matches the oracle's term-by-term
    sum to the parity bar."""
    print(f"case split D={D} L={L}", flush=True)
    b = E.SG4Basis(D, L)
    o = O.SG4(D, L)
    rng = np.random.default_rng(99 + D)
    V = rng.standard_normal(o.sum_nq)
    ck = -0.5 * ck_scale * (1.0 + 0.2 * rng.standard_normal(D))
    op = E.SG4Operator(b, max_nvec=8, potential=V, ck=ck)
    split, nnz = op.split_info()
    assert split and nnz >= b.nb_ba
    psi = rng.standard_normal((b.nb_ba, 6))
    ref = o.hpsi(psi, V, ck)
    for prec, tol in ((E.PREC_FP64, TOL), (E.PREC_F32, 1e-5)):
        got = op.apply_host(psi, prec)
        print(f"  prec {prec}: max rel {rel(got, ref):.2e}, nnz {nnz}")
        assert rel(got, ref) < tol


def test_split_path_not_taken_for_rectangular_bases(cuda):
    """nq > nb (grid level above the basis level): the terms are not square,
    the plan keeps the grid path and still matches the oracle."""
    o = O.SG4(3, 4, nq_A=3, nb_A=2)
    b = E.SG4Basis(3, 4, modes=[dict(nq_A=3, nb_A=2)] * 3)
    V = np.random.default_rng(5).standard_normal(o.sum_nq)
    op = E.SG4Operator(b, max_nvec=4, potential=V, ck=-0.5 * np.ones(3))
    assert not op.split_info()[0]
    psi = np.random.default_rng(6).standard_normal((b.nb_ba, 3))
    assert rel(op.apply_host(psi), o.hpsi(psi, V, -0.5 * np.ones(3))) < TOL


def test_jacobi_device_grids_equal_host_grids(cuda):
    """V, G, Jac evaluated on the device (sg4_plan_set_jacobi_triatomic) give
    the same type-10 operator as the host-evaluated grids."""
    import elvibrot_amd.potentials as P
    op_dev = P.jacobi_triatomic_operator(6, max_nvec=4)
    op_host = P.jacobi_triatomic_operator(6, max_nvec=4, device_grids=False)
    psi = np.random.default_rng(8).standard_normal((op_dev.basis.nb_ba, 3))
    assert rel(op_dev.apply_host(psi), op_host.apply_host(psi)) < 1e-14


def test_launch_configurations_keep_parity(cuda):
    """Every item order (plain, per term, classes of terms on one XCD) and the
    768-thread FP32 kernel give the same operator; out-of-range launch
    parameters are rejected."""
    D, L, nvec = 6, 6, 16
    b, op, o, V = hh_pair(D, L)
    psi = np.random.default_rng(31).standard_normal((b.nb_ba, nvec))
    ref = o.hpsi(psi, V, -0.5 * np.ones(D))
    for xcd in (0, 1, 2, 3, 4):
        op.set_config(nvec, E.PREC_FP64, 512, 163840, 1024, xcd)
        assert rel(op.apply_host(psi), ref) < TOL, xcd
    op.set_config(nvec, E.PREC_F32, 768, 163840, 1024, 2)
    assert rel(op.apply_host(psi, E.PREC_F32), ref) < 1e-5
    for bad in ((768, 65536, 0, 0, E.PREC_FP64), (512, 200000, 0, 0, E.PREC_FP64),
                (512, 65536, -1, 0, E.PREC_FP64), (512, 65536, 0, 99, E.PREC_FP64)):
        with pytest.raises(E.SG4Error):
            op.set_config(nvec, bad[4], *bad[:4])


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_F32, E.PREC_F32X2])
def test_block_wider_than_workgroup(cuda, prec):
    """1024 vectors in one block: the one-point term's chunk holds more vectors
    than a workgroup has threads, and every vector must still get its term."""
    D, L, nvec = 6, 3, 1024
    b, op, o, V = hh_pair(D, L, nvec)
    psi = np.random.default_rng(77).standard_normal((b.nb_ba, nvec))
    ref = o.hpsi(psi, V, -0.5 * np.ones(D))
    got = op.apply_host(psi, prec)
    err = rel(got, ref)
    print(f"6D L=3 x1024 prec {prec}: {err:.2e}")
    assert err < (TOL if prec == E.PREC_FP64 else 5e-6)


def test_graph_recaptured_after_plan_changes(cuda):
    """A captured graph points at plan buffers and was captured with one launch
    configuration: after the grouping, the potential or the configuration
    changes, sg4_apply_graph must give what sg4_apply_rows gives."""
    import torch
    D, L, nvec = 5, 4, 4
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=nvec)
    X = torch.randn(b.nb_ba, nvec, device=cuda, dtype=torch.float64,
                    generator=torch.Generator(device=cuda).manual_seed(3))
    Yg, Yr = torch.empty_like(X), torch.empty_like(X)

    def same(prec=E.PREC_FP64, tol=1e-13):
        op.apply_graph(X, Yg, prec)
        op.apply_rows(X, Yr, prec)
        torch.cuda.synchronize()
        err = (torch.max(torch.abs(Yg - Yr)) / torch.max(torch.abs(Yr))).item()
        assert err < tol, err

    same()
    op.set_grouping(False)
    same()
    op.set_grouping(True)
    same()
    E._capi.check(E.lib().sg4_plan_set_henon_heiles(op._h, 0.2))
    same()
    op.set_config(nvec, E.PREC_FP64, nthreads=256, lds_budget=32768, grid=0, xcd=1)
    same()
    # FP32 replay after a field change reads the refreshed FP32 potential
    mu = torch.ones(b.sum_nq, device=cuda, dtype=torch.float64)
    same(E.PREC_F32, 1e-6)
    op.set_field(mu, 0.3)
    same(E.PREC_F32, 1e-6)
    op.set_field(None, 0.0)


def test_deterministic_sum_bitwise_across_launches(cuda):
    """FP64 split path with the deterministic sum: bitwise identical results
    for repeated applies and for different launch configurations (item order,
    LDS budget, persistent grid), equal to the default FP64-atomic apply to
    rounding."""
    import torch
    D, L, nvec = 6, 6, 16
    b = E.SG4Basis(D, L)
    op = E.SG4Operator(b, max_nvec=nvec)
    X = torch.randn(b.nb_ba, nvec, device=cuda, dtype=torch.float64,
                    generator=torch.Generator(device=cuda).manual_seed(13))
    Yd = torch.empty_like(X)
    op.apply_rows(X, Yd)
    op.set_deterministic(True)
    outs = []
    for cfg in [(256, 65536, 0, 0), (512, 163840, 2048, 3), (512, 32768, 1024, 1), (256, 81920, 0, 2)]:
        op.set_config(nvec, E.PREC_FP64, *cfg)
        Y = torch.empty_like(X)
        op.apply_rows(X, Y)
        outs.append(Y)
    torch.cuda.synchronize()
    assert all(torch.equal(outs[0], y) for y in outs[1:])
    assert (torch.max(torch.abs(outs[0] - Yd)) / torch.max(torch.abs(Yd))).item() < 1e-13
