"""Wave-packet files in ElVibRot's formats (SRC/sub_WP/sub_module_psi_io.f90).

The packed-row index tuples written to the files are checked against the
oracle's packed basis (lex order of the reference's nDindB); the line format
against the version-0 wave-packet lines that the reference's own test input
Working_tests/exa_hcn-test/dat_2DSG_type4:71-74 holds (fixture
tests/golden/psi_v0_2d_dat_2DSG_type4.txt); the rest by round trips."""
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
