"""Wave-packet files in ElVibRot's formats (SRC/sub_WP/sub_module_psi_io.f90).

The packed-row index tuples written to the files are checked against the
oracle's packed basis (lex order of the reference's nDindB); the line format
against the version-0 wave-packet lines that the reference's own test input
Working_tests/exa_hcn-test/dat_2DSG_type4:71-74 holds (fixture
tests/golden/psi_v0_2d_dat_2DSG_type4.txt) and the other inline packets of
the exa_hcn inputs; version 1 and the restart file both ways through a
Fortran runtime (tests/fortran_io/read_wp_files.f90 built with amdflang:
files written here read with the reference writer's statements, files that
runtime writes read here); the rest by round trips."""
import os
import struct

import numpy as np
import pytest

import elvibrot_amd as E
from elvibrot_amd import psi_io
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("D,L,LB", [(2, 6, None), (4, 4, None), (6, 3, None), (4, 5, 3)])
def test_packed_indices_match_oracle(D, L, LB):
    b = E.SG4Basis(D, L, LB)
    o = O.SG4(D, L, LB)
    assert np.array_equal(psi_io.packed_indices(b), o.packed)


def test_reference_version0_lines(tmp_path):
    """The two one-line packets of the reference input (indices 1 1 and 2 1,
    i_h = i_e = 1, coefficient 1.0, each closed by 'end wp0'), behind a
    one-integer header as Read_header_saveFile_psi accepts."""
    b = E.SG4Basis(2, 4)
    body = open(os.path.join(HERE, "golden", "psi_v0_2d_dat_2DSG_type4.txt")).read()
    p = tmp_path / "wp.txt"
    p.write_text(" 2\n" + body)
    psi, info = psi_io.read_psi(str(p), b)
    assert info["version"] == 0 and info["nb_psi"] == 2 and info["skipped"] == 0
    idx = psi_io.packed_indices(b)
    r0 = int(np.nonzero((idx == [0, 0]).all(1))[0][0])
    r1 = int(np.nonzero((idx == [1, 0]).all(1))[0][0])
    want = np.zeros((b.nb_ba, 2))
    want[r0, 0] = 1.0
    want[r1, 1] = 1.0
    assert np.array_equal(psi, want)


def test_version0_layout_and_round_trip(tmp_path):
    b = E.SG4Basis(3, 3)
    rng = np.random.default_rng(1)
    psi = rng.standard_normal((b.nb_ba, 3))
    psi[5, 1] = 1e-12
    p = tmp_path / "v0.txt"
    psi_io.write_psi(str(p), psi, b, version=0, epsi=1e-10)
    lines = p.read_text().splitlines()
    assert lines[0].split() == ["3", str(b.nb_ba)]
    data = [l for l in lines[1:] if "end" not in l]
    assert all(len(l.split()) == 3 + 3 for l in data)      # D indices, i_h, i_e, a
    assert [l.split()[:2] for l in lines if "end" in l] == [["end", "wp"]] * 3
    assert len(data) == 3 * b.nb_ba - 1                      # |a| < epsi dropped
    got, info = psi_io.read_psi(str(p), b)
    want = psi.copy()
    want[5, 1] = 0.0
    assert np.array_equal(got, want)


@pytest.mark.parametrize("formatted", [True, False])
def test_version1_round_trip(tmp_path, formatted):
    b = E.SG4Basis(4, 3)
    rng = np.random.default_rng(2)
    psi = rng.standard_normal((b.nb_ba, 2)) + 1j * rng.standard_normal((b.nb_ba, 2))
    p = tmp_path / "v1.dat"
    psi_io.write_psi(str(p), psi, b, version=1, T=12.5, formatted=formatted)
    got, info = psi_io.read_psi(str(p), b, cplx=True, formatted=formatted)
    assert info["version"] == 1 and info["nb_psi"] == 2 and info["T"] == [12.5, 12.5]
    assert np.array_equal(got, psi)
    if not formatted:
        raw = p.read_bytes()
        assert struct.unpack_from("<i", raw, 0)[0] == 1 and raw[4:5] == b"X"


def test_read_into_smaller_basis_skips_absent_functions(tmp_path):
    big, small = E.SG4Basis(3, 5), E.SG4Basis(3, 3)
    psi = np.random.default_rng(3).standard_normal((big.nb_ba, 1))
    p = tmp_path / "big.txt"
    psi_io.write_psi(str(p), psi, big)
    got, info = psi_io.read_psi(str(p), small)
    assert info["skipped"] == big.nb_ba - small.nb_ba
    rows = {tuple(r): n for n, r in enumerate(psi_io.packed_indices(big))}
    sel = [rows[tuple(r)] for r in psi_io.packed_indices(small)]
    assert np.array_equal(got[:, 0], psi[sel, 0])


def test_version1_complex_is_written_as_list_directed_complex_constants(tmp_path):
    """Write_Psi_nDBasis version 1 writes ``T, CvecB(:)`` list-directed
    (sub_module_psi_io.f90:1621): every coefficient is a constant (re,im)."""
    b = E.SG4Basis(2, 2)
    psi = (np.arange(b.nb_ba) + 1j * (np.arange(b.nb_ba) + 0.5))[:, None]
    p = tmp_path / "v1c.dat"
    psi_io.write_psi(str(p), psi, b, version=1, T=2.0)
    body = p.read_text().splitlines()[5 + b.nb_ba:]
    assert float(body[0]) == 2.0
    toks = " ".join(body[1:]).split()
    assert len(toks) == b.nb_ba and all(t.startswith("(") and t.endswith(")") for t in toks)
    re_, im_ = toks[3][1:-1].split(",")
    assert float(re_) == 3.0 and float(im_) == 3.5


def test_version1_reader_accepts_fortran_list_directed_forms(tmp_path):
    """Fortran list-directed records: blanks and commas, D exponents, repeat
    counts r*c, values over several records; and the (re, im) real pairs of
    files written before the complex constants."""
    b = E.SG4Basis(2, 2)
    idx = psi_io.packed_indices(b) + 1
    head = " &HEADERFILE\n VERSION_FILE=1,\n NB_PSI=1,\n NB_TOT=%d,\n /\n" % b.nb_ba
    lst = "".join(" %d %d 1 1\n" % tuple(r) for r in idx)
    n = b.nb_ba
    vals = ["(%d.0D0 , -1.0d-1)" % k for k in range(n - 2)] + ["2*(0.0,0.0)"]
    text = head + lst + "   3.0000000000000000     ,\n" + "\n".join(
        ", ".join(vals[a:a + 3]) for a in range(0, len(vals), 3)) + "\n"
    p = tmp_path / "fortran_v1.dat"
    p.write_text(text)
    got, info = psi_io.read_psi(str(p), b, cplx=True)
    want = np.array([k - 0.1j for k in range(n - 2)] + [0, 0])
    assert info["T"] == [3.0] and np.array_equal(got[:, 0], want)
    # pairs of reals
    pairs = " ".join("%d.0 -0.1" % k for k in range(n - 2)) + " 0.0 0.0 0.0 0.0"
    p.write_text(head + lst + " 3.0 " + pairs + "\n")
    got, _ = psi_io.read_psi(str(p), b, cplx=True)
    assert np.array_equal(got[:, 0], want)


def test_restart_round_trip_and_fortran_record_layout(tmp_path):
    """SaveWP_restart / ReadWP_restart (sub_module_propagation.f90:379-495):
    record ``T nb_WP nb_tot``, then each packet's complex coefficients from a
    new record on."""
    rng = np.random.default_rng(5)
    psis = rng.standard_normal((3, 7)) + 1j * rng.standard_normal((3, 7))
    p = tmp_path / "file_WP_restart"
    psi_io.write_restart(str(p), 1.25, psis)
    lines = p.read_text().splitlines()
    head = lines[0].split()
    assert float(head[0]) == 1.25 and head[1:] == ["3", "7"]
    assert len(lines) == 1 + 3 * 4        # 7 constants, two per record, per packet
    T, got = psi_io.read_restart(str(p), 3, 7)
    assert T == 1.25 and np.array_equal(got, psis)


def test_restart_reader_fortran_text_missing_time_and_size_checks(tmp_path):
    p = tmp_path / "r"
    # gfortran-style list-directed output, a packet over two records, a repeat count
    p.write_text("   0.50000000000000000                2           3\n"
                 " (  1.0000000000000000     , -2.0000000000000000     ) (0.5D0,0.25D0)\n"
                 " (0.0,1.0)\n"
                 " 3*(  0.0000000000000000     ,  0.0000000000000000     )\n")
    T, got = psi_io.read_restart(str(p), 2, 3)
    assert T == 0.5
    assert np.array_equal(got, np.array([[1 - 2j, 0.5 + 0.25j, 1j], [0, 0, 0]]))
    with pytest.raises(ValueError, match="inconsistent"):
        psi_io.read_restart(str(p), 2, 4)
    p.write_text(" (1.0,0.0) (2.0,0.0)\n")    # no T record: no restart, T0 = 0
    assert psi_io.read_restart(str(p), 1, 2) == (0.0, None)
    p.write_text(" (1.0,0.0) (2.0,0.0)\n (3.0,0.0)\n")
    assert psi_io.read_restart(str(p), 1, 2) == (0.0, None)


def test_restart_header_over_several_records(tmp_path):
    """``read(nioWP,*) T, nb_WP, nb_tot`` is one list-directed read: the three
    values may sit on separate records, and the packets start on the record
    after the last of them."""
    p = tmp_path / "r"
    p.write_text(" 0.25D0\n 1,\n 2\n (1.0,2.0) (3.0,-4.0)\n")
    T, got = psi_io.read_restart(str(p), 1, 2)
    assert T == 0.25 and np.array_equal(got, np.array([[1 + 2j, 3 - 4j]]))
    p.write_text(" 0.25 1\n 2 \n (5.0,6.0)\n (7.0,8.0)\n")
    T, got = psi_io.read_restart(str(p), 1, 2)
    assert T == 0.25 and np.array_equal(got, np.array([[5 + 6j, 7 + 8j]]))


def test_propagation_writes_restart_every_step_and_resumes(tmp_path):
    """propagate_field's restart file holds the state and time after the last
    written step; resuming from it with the remaining segments gives the same
    final state as the uninterrupted run (a host-side stand-in propagator:
    the file logic does not touch the GPU)."""
    import torch
    from elvibrot_amd.propagation import propagate_field

    class Stub:
        dt = 0.1

        def __init__(self):
            self.e = 0.0

        def set_field(self, e):
            self.e = e

        def step(self, psi, npoly=None):
            return psi * complex(np.cos(1 + self.e), -np.sin(1 + self.e)) + self.e, 3

    psi0 = torch.tensor(np.random.default_rng(6).standard_normal((2, 5)) + 0j)
    eps = [0.1, -0.2, 0.3]
    p = tmp_path / "file_WP_restart"
    full, _, _ = propagate_field(Stub(), psi0, eps, 2, restart_path=str(p), restart_every=4)
    T, last = psi_io.read_restart(str(p), 2, 5)
    assert abs(T - 0.6) < 1e-15 and np.array_equal(last, full.numpy())
    part, _, _ = propagate_field(Stub(), psi0, eps[:2], 2, restart_path=str(p))
    T, mid = psi_io.read_restart(str(p), 2, 5)
    assert abs(T - 0.4) < 1e-15
    rest, _, _ = propagate_field(Stub(), torch.tensor(mid), eps[2:], 2, restart_path=str(p), t0=T)
    assert np.array_equal(rest.numpy(), full.numpy())
    assert abs(psi_io.read_restart(str(p), 2, 5)[0] - 0.6) < 1e-15


@pytest.mark.parametrize("name,D,L,tuples", [
    ("psi_v0_1d_dat_Davidson_1D_WithGrid.txt", 1, 4, [(1,), (2,)]),
    ("psi_v0_2d_dat_Davidson_AutoContract.txt", 2, 4, [(1, 1), (1, 2)]),
    ("psi_v0_3d_dat_Davidson.txt", 3, 4, [(1, 1, 1), (2, 1, 1), (1, 2, 1), (1, 1, 2)]),
    ("psi_v0_3d_dat_Davidson_WithGrid.txt", 3, 4,
     [(1, 1, 1), (4, 1, 1), (6, 1, 1), (1, 1, 2), (1, 2, 1)]),
])
def test_reference_inline_wave_packets(tmp_path, name, D, L, tuples):
    """The initial packets of the reference's exa_hcn inputs (version-0 lines
    'i_1 .. i_D i_h i_e a', closed by 'end wp0', one with a leading blank):
    each is the unit vector on the packed row of its 1-based index tuple."""
    b = E.SG4Basis(D, L)
    body = open(os.path.join(HERE, "golden", name)).read()
    p = tmp_path / "wp.txt"
    p.write_text(f" {len(tuples)}\n" + body)
    psi, info = psi_io.read_psi(str(p), b)
    assert info["nb_psi"] == len(tuples) and info["skipped"] == 0
    idx = psi_io.packed_indices(b)
    want = np.zeros((b.nb_ba, len(tuples)))
    for j, tup in enumerate(tuples):
        want[int(np.nonzero((idx == np.array(tup) - 1).all(1))[0][0]), j] = 1.0
    assert np.array_equal(psi, want)


@pytest.fixture(scope="module")
def fortran_reader(tmp_path_factory):
    """tests/fortran_io/read_wp_files.f90 built with amdflang."""
    exe = tmp_path_factory.mktemp("fio") / "read_wp_files"
    src = os.path.join(HERE, "fortran_io", "read_wp_files.f90")
    r = __import__("subprocess").run(["/opt/rocm/bin/amdflang", "-O1", src, "-o", str(exe)],
                                     capture_output=True, text=True, cwd=str(exe.parent))
    assert r.returncode == 0, r.stderr
    return str(exe)


def _run_reader(exe, mode, D, cplx, fin, tmp_path):
    out = tmp_path / (mode + ".out")
    r = __import__("subprocess").run([exe, mode, str(D), str(int(cplx)), str(fin), str(out)],
                                     capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    nb_psi, nb_tot, version = np.frombuffer(raw[:12], dtype=np.int32)
    pos = 12
    ind = None
    if mode != "rst":
        ind = np.frombuffer(raw[pos:pos + 4 * (D + 2) * nb_tot], dtype=np.int32).reshape(nb_tot, D + 2)
        pos += 4 * (D + 2) * nb_tot
    per = 1 + (2 if cplx else 1) * nb_tot
    vals = np.frombuffer(raw[pos:], dtype=np.float64).reshape(nb_psi, per)
    T = vals[:, 0]
    v = vals[:, 1:]
    if cplx:
        v = v[:, 0::2] + 1j * v[:, 1::2]
    return int(nb_psi), int(nb_tot), int(version), ind, T, v


@pytest.mark.parametrize("formatted", [True, False])
@pytest.mark.parametrize("cplx", [False, True])
def test_version1_files_read_by_fortran_runtime(tmp_path, fortran_reader, formatted, cplx):
    """Version-1 files written here, read back by amdflang's runtime with the
    statements of the reference's writer (namelist &headerFile, list-directed
    index lines and 'T, vec(:)'; or the unformatted sequential records):
    every count, index tuple, T and coefficient comes back bit for bit.
    (The reference's own reader, Read_psi_nDBasis case 1,
    sub_module_psi_io.f90:797-812, reads the vector without the T its writer
    puts first, :1606-1617; these files follow the writer.)"""
    b = E.SG4Basis(3, 4)
    rng = np.random.default_rng(11)
    psi = rng.standard_normal((b.nb_ba, 3))
    if cplx:
        psi = psi + 1j * rng.standard_normal((b.nb_ba, 3))
    p = tmp_path / "wp1"
    psi_io.write_psi(str(p), psi, b, version=1, T=0.75, formatted=formatted)
    nb_psi, nb_tot, version, ind, T, v = _run_reader(
        fortran_reader, "v1f" if formatted else "v1u", 3, cplx, p, tmp_path)
    assert (nb_psi, nb_tot, version) == (3, b.nb_ba, 1)
    assert np.array_equal(ind[:, :3], psi_io.packed_indices(b) + 1)
    assert np.all(ind[:, 3:] == 1)
    assert np.all(T == 0.75)
    assert np.array_equal(v, psi.T)


def test_restart_file_read_by_fortran_runtime(tmp_path, fortran_reader):
    """The propagation restart file read by amdflang's list-directed input as
    ReadWP_restart reads it: T, nb_WP, nb_tot, then every packet's complex
    coefficients, bit for bit."""
    rng = np.random.default_rng(12)
    psis = rng.standard_normal((4, 37)) + 1j * rng.standard_normal((4, 37))
    p = tmp_path / "file_WP_restart"
    psi_io.write_restart(str(p), 12.5, psis)
    nb_psi, nb_tot, _, _, T, v = _run_reader(fortran_reader, "rst", 0, True, p, tmp_path)
    assert (nb_psi, nb_tot) == (4, 37)
    assert np.all(T == 12.5)
    assert np.array_equal(v, psis)


def _dump(path, D, psi, T, ind=None):
    nb, n = psi.shape
    parts = [np.array([n, nb, 1], dtype=np.int32).tobytes()]
    if ind is not None:
        parts.append(np.ascontiguousarray(ind, dtype=np.int32).tobytes())
    for j in range(n):
        v = psi[:, j]
        if np.iscomplexobj(v):
            v = np.column_stack([v.real, v.imag]).ravel()
        parts.append(np.concatenate([[T], v]).astype(np.float64).tobytes())
    path.write_bytes(b"".join(parts))


@pytest.mark.parametrize("cplx", [False, True])
def test_version1_file_written_by_fortran_runtime_is_read(tmp_path, fortran_reader, cplx):
    """amdflang writes a version-1 file with the reference writer's statements
    (namelist, list-directed index lines, 'T, vec(:)' in the runtime's own
    output form); read_psi returns every coefficient and T bit for bit."""
    b = E.SG4Basis(3, 3)
    rng = np.random.default_rng(13)
    psi = rng.standard_normal((b.nb_ba, 2)) * 10.0 ** rng.integers(-30, 30, (b.nb_ba, 2))
    if cplx:
        psi = psi + 1j * rng.standard_normal((b.nb_ba, 2))
    ind = np.hstack([psi_io.packed_indices(b) + 1, np.ones((b.nb_ba, 2), dtype=np.int32)])
    d, f = tmp_path / "dump", tmp_path / "wp_fortran"
    _dump(d, 3, psi, -3.25, ind)
    r = __import__("subprocess").run([fortran_reader, "w1f", "3", str(int(cplx)), str(d), str(f)],
                                     capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    got, info = psi_io.read_psi(str(f), b, cplx=cplx)
    assert info["version"] == 1 and info["nb_psi"] == 2 and info["T"] == [-3.25, -3.25]
    assert np.array_equal(got, psi)


def test_restart_file_written_by_fortran_runtime_is_read(tmp_path, fortran_reader):
    rng = np.random.default_rng(14)
    psis = rng.standard_normal((3, 11)) + 1j * rng.standard_normal((3, 11))
    d, f = tmp_path / "dump", tmp_path / "file_WP_restart"
    _dump(d, 0, psis.T, 7.5)
    r = __import__("subprocess").run([fortran_reader, "wrst", "0", "1", str(d), str(f)],
                                     capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    T, got = psi_io.read_restart(str(f), 3, 11)
    assert T == 7.5 and np.array_equal(got, psis)


@pytest.mark.parametrize("version,formatted", [(0, True), (1, True), (1, False)])
def test_two_state_files_round_trip(tmp_path, version, formatted):
    """nb_e = 2: the components one after the other, i_e the state index
    (the reference writer's i_e, i_h, i_b loop order), back bit for bit; a
    one-state read keeps the first state and skips the second's rows."""
    b = E.SG4Basis(3, 3)
    nb = b.nb_ba
    rng = np.random.default_rng(21)
    psi = rng.standard_normal((2 * nb, 2)) + 1j * rng.standard_normal((2 * nb, 2))
    p = tmp_path / f"wp2_{version}_{formatted}"
    psi_io.write_psi(str(p), psi, b, version=version, formatted=formatted, nb_e=2)
    got, info = psi_io.read_psi(str(p), b, cplx=True, formatted=formatted, nb_e=2)
    assert np.array_equal(got, psi) and info["skipped"] == 0 and info["nb_tot"] == 2 * nb
    one, info1 = psi_io.read_psi(str(p), b, cplx=True, formatted=formatted)
    assert np.array_equal(one, psi[:nb])
    assert info1["skipped"] == 2 * nb            # the second state's rows of both vectors
    if formatted and version == 0:
        rows = [l.split() for l in open(p).read().splitlines()[1:] if "end" not in l]
        ie = np.array([int(r[b.D + 1]) for r in rows[:2 * nb]])
        assert np.array_equal(ie, np.repeat([1, 2], nb))
    with pytest.raises(ValueError):
        psi_io.write_psi(str(p), psi[:nb + 1], b, nb_e=2)


def test_two_state_version1_read_by_fortran_runtime(tmp_path, fortran_reader):
    """A two-state version-1 file read by amdflang's runtime: nb_tot = 2 nb_ba
    index lines with i_e = 1 then 2, and every coefficient bit for bit."""
    b = E.SG4Basis(3, 3)
    nb = b.nb_ba
    psi = np.random.default_rng(22).standard_normal((2 * nb, 2))
    p = tmp_path / "wp2"
    psi_io.write_psi(str(p), psi, b, version=1, T=1.5, nb_e=2)
    nb_psi, nb_tot, version, ind, T, v = _run_reader(fortran_reader, "v1f", 3, False, p, tmp_path)
    assert (nb_psi, nb_tot, version) == (2, 2 * nb, 1)
    assert np.array_equal(ind[:, 3], np.ones(2 * nb)) and np.array_equal(ind[:, 4], np.repeat([1, 2], nb))
    assert np.array_equal(ind[:nb, :3], ind[nb:, :3])
    assert np.all(T == 1.5) and np.array_equal(v, psi.T)
