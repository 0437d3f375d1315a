"""C ABI library: loads, exports every symbol of include/sg4hpsi.h, and its
host-side set-up reproduces the oracle's tables (CPU only, no compute calls
on a device)."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
import elvibrot_amd as E
from elvibrot_amd import _capi
from oracle import oracle as O


def header_functions():
    src = open(os.path.join(ROOT, "include", "sg4hpsi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sg4_[A-Za-z0-9_]+)\s*\(", src)))


def test_header_matches_binding_list():
    assert header_functions() == sorted(_capi.EXPORTS)


def test_library_exports_every_header_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = [f for f in header_functions() if f not in syms]
    assert not missing, missing
    L = E.lib()
    for f in header_functions():
        assert getattr(L, f) is not None


@pytest.mark.parametrize("D,L", [(6, 3), (6, 6), (21, 2), (4, 5)])
def test_setup_tables_equal_oracle(D, L):
    b = E.SG4Basis(D, L)
    o = O.SG4(D, L)
    assert (b.nterms, b.nb_ba, b.sum_nb, b.sum_nq) == (o.nterms, o.nb_ba, o.sum_nb, o.sum_nq)
    assert np.array_equal(b.levels(), o.lev)
    assert np.array_equal(b.weights(), o.weight)
    assert np.array_equal(b.map(), o.map)
    for k in range(D):
        for l in range(L + 1):
            bb, ob = b.basis1d(k, l), o.basis1d(k, l)
            for key in ("x", "w", "d0", "d1", "d2"):
                assert np.max(np.abs(bb[key] - ob[key])) <= 1e-14 * max(1, np.abs(ob[key]).max())
    t = o.nterms // 2
    assert np.array_equal(b.term_grid(t), o.term_grid(t))


def test_setup_lb_smaller_than_lg_has_absent_entries():
    b = E.SG4Basis(3, 4, LB=2)
    o = O.SG4(3, 4, LB=2)
    assert b.nb_ba == o.nb_ba
    assert np.array_equal(b.map(), o.map)
    assert (b.map() == -1).any()


def test_setup_rejects_bad_input():
    with pytest.raises(E.SG4Error):
        E.SG4Basis(0, 3)
    with pytest.raises(E.SG4Error):
        E.SG4Basis(3, 2, LB=5)
    with pytest.raises(E.SG4Error):
        E.SG4Basis(2, 3, modes=[dict(nq_A=1, nq_B=1, nb_A=2, nb_B=1)] * 2)


def test_setup_jacobi_bases_equal_oracle():
    import elvibrot_amd.potentials as P
    b = P.jacobi_triatomic_basis(6)
    sc = np.sqrt(P.AMU * np.sqrt(0.4 / P.AMU))
    o = O.SG4(3, 6, kind=[0, 0, 1], Q0=[3.0, 2.0, 0.0], scaleQ=[sc, sc, 1.0], nq_A=2, nb_A=2)
    assert np.array_equal(b.map(), o.map)
    for k in range(3):
        for l in range(7):
            bb, ob = b.basis1d(k, l), o.basis1d(k, l)
            for key in ("x", "w", "d0", "d1", "d2"):
                assert np.max(np.abs(bb[key] - ob[key])) <= 1e-13 * max(1, np.abs(ob[key]).max())
