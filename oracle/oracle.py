"""ctypes binding of the CPU oracle (``sg4_oracle.c``).

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``, always as the checker, never as
the thing measured or shipped.  The product (``elvibrot_amd``) never imports
it.

The oracle restates ElVibRot's SG4 set-up and FP64 H.psi
(SRC/sub_Operator/sub_OpPsi_SG4.f90:681-1666,
SRC/sub_Basis/sub_Basis_SG4/sub_module_basis_BtoG_GtoB_SG4.f90,
SRC/sub_Basis/sub_quadra_SparseBasis.f90:922-1434); see sg4_oracle.c for the
per-function citations and pins.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc_sg4.so")
_lib = None

D_P = C.POINTER(C.c_double)
I_P = C.POINTER(C.c_int)
L_P = C.POINTER(C.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_gauss_hermite.argtypes = [C.c_int, D_P, D_P]
        L.orc_gauss_legendre.argtypes = [C.c_int, D_P, D_P]
        L.orc_ho_basis.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double] + [D_P] * 5
        L.orc_legendre_basis.argtypes = [C.c_int, C.c_int] + [D_P] * 5
        L.orc_grid_deriv.argtypes = [C.c_int, C.c_int] + [D_P] * 6
        L.orc_sg_create.restype = C.c_void_p
        L.orc_sg_create.argtypes = [C.c_int, C.c_int, C.c_int, I_P, I_P, I_P, I_P, I_P, D_P, D_P]
        L.orc_sg_destroy.argtypes = [C.c_void_p]
        for f in ("orc_sg_nterms", "orc_sg_nb_ba", "orc_sg_sum_nb", "orc_sg_sum_nq"):
            getattr(L, f).restype = C.c_int64
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_sg_tables.argtypes = [C.c_void_p, I_P, D_P, L_P, L_P, I_P, I_P]
        L.orc_sg_basis1d.argtypes = [C.c_void_p, C.c_int, C.c_int, I_P, I_P] + [D_P] * 7
        L.orc_sg_term_grid.argtypes = [C.c_void_p, C.c_int64, D_P]
        L.orc_henon_heiles.restype = C.c_double
        L.orc_henon_heiles.argtypes = [C.c_int, D_P, C.c_double]
        L.orc_sg_grid_hh.argtypes = [C.c_void_p, C.c_double, D_P]
        L.orc_hpsi_range.argtypes = [C.c_void_p, C.c_int, D_P, D_P, D_P, D_P, D_P, C.c_int,
                                     D_P, D_P, C.c_int64, C.c_int64, C.c_int64]
        L.orc_hpsi_absbound.argtypes = [C.c_void_p, C.c_int, D_P, D_P, D_P, D_P, D_P, C.c_int,
                                        D_P, D_P]
        L.orc_hpsi_ld.argtypes = [C.c_void_p, D_P, D_P, C.c_int, D_P, D_P]
        L.orc_hpsi_t1g.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), D_P, C.c_int, D_P, D_P]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(D_P)


def _i(a):
    return None if a is None else a.ctypes.data_as(I_P)


def _l(a):
    return None if a is None else a.ctypes.data_as(L_P)


def gauss_hermite(n):
    x = np.zeros(n)
    w = np.zeros(n)
    assert lib().orc_gauss_hermite(n, _d(x), _d(w)) == 0
    return x, w


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    assert lib().orc_gauss_legendre(n, _d(x), _d(w)) == 0
    return x, w


def ho_basis(nq, nb, Q0=0.0, scaleQ=1.0):
    x, w = np.zeros(nq), np.zeros(nq)
    d0, d1, d2 = (np.zeros((nb, nq)) for _ in range(3))  # transposed view = (nq, nb) F-order
    assert lib().orc_ho_basis(nq, nb, Q0, scaleQ, _d(x), _d(w), _d(d0), _d(d1), _d(d2)) == 0
    return x, w, d0.T, d1.T, d2.T


def legendre_basis(nq, nb):
    x, w = np.zeros(nq), np.zeros(nq)
    d0, d1, d2 = (np.zeros((nb, nq)) for _ in range(3))
    assert lib().orc_legendre_basis(nq, nb, _d(x), _d(w), _d(d0), _d(d1), _d(d2)) == 0
    return x, w, d0.T, d1.T, d2.T


def grid_deriv(d0, d1, d2):
    nq, nb = d0.shape
    f = lambda a: np.asfortranarray(a)
    d0f, d1f, d2f = f(d0), f(d1), f(d2)
    g0, g1, g2 = (np.zeros((nq, nq), order="F") for _ in range(3))
    assert lib().orc_grid_deriv(nq, nb, _d(d0f), _d(d1f), _d(d2f), _d(g0), _d(g1), _d(g2)) == 0
    return g0, g1, g2


class SG4:
    """Oracle SG4 basis: term set, weights, packed basis, map, 1-D bases."""

    def __init__(self, D, L, LB=None, kind=None, nq_A=1, nq_B=2, nb_A=None, nb_B=None,
                 Q0=None, scaleQ=None):
        LB = L if LB is None else LB
        vec = lambda v, dt: np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=dt), (D,)))
        self.D, self.L, self.LB = D, L, LB
        self.kind = vec(0 if kind is None else kind, np.int32)
        self.nqA, self.nqB = vec(nq_A, np.int32), vec(nq_B, np.int32)
        self.nbA = vec(nq_A if nb_A is None else nb_A, np.int32)
        self.nbB = vec(nq_B if nb_B is None else nb_B, np.int32)
        self.Q0 = vec(0.0 if Q0 is None else Q0, np.float64)
        self.scaleQ = vec(1.0 if scaleQ is None else scaleQ, np.float64)
        self._h = lib().orc_sg_create(D, L, LB, _i(self.kind), _i(self.nqA), _i(self.nqB),
                                      _i(self.nbA), _i(self.nbB), _d(self.Q0), _d(self.scaleQ))
        if not self._h:
            raise ValueError("oracle SG4 set-up failed")
        L_ = lib()
        self.nterms = L_.orc_sg_nterms(self._h)
        self.nb_ba = L_.orc_sg_nb_ba(self._h)
        self.sum_nb = L_.orc_sg_sum_nb(self._h)
        self.sum_nq = L_.orc_sg_sum_nq(self._h)
        self.lev = np.zeros((self.nterms, D), dtype=np.int32)
        self.weight = np.zeros(self.nterms)
        self.boff = np.zeros(self.nterms + 1, dtype=np.int64)
        self.qoff = np.zeros(self.nterms + 1, dtype=np.int64)
        self.map = np.zeros(self.sum_nb, dtype=np.int32)
        self.packed = np.zeros((self.nb_ba, D), dtype=np.int32)
        L_.orc_sg_tables(self._h, _i(self.lev), _d(self.weight), _l(self.boff), _l(self.qoff),
                         _i(self.map), _i(self.packed))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_sg_destroy(self._h)
            self._h = None

    def basis1d(self, k, l):
        nq, nb = C.c_int(), C.c_int()
        lib().orc_sg_basis1d(self._h, k, l, C.byref(nq), C.byref(nb), *([None] * 7))
        nq, nb = nq.value, nb.value
        x, w = np.zeros(nq), np.zeros(nq)
        d0, d1, d2 = (np.zeros((nb, nq)) for _ in range(3))
        g1, g2 = np.zeros((nq, nq)), np.zeros((nq, nq))
        lib().orc_sg_basis1d(self._h, k, l, C.byref(C.c_int()), C.byref(C.c_int()), _d(x), _d(w),
                             _d(d0), _d(d1), _d(d2), _d(g1), _d(g2))
        return dict(nq=nq, nb=nb, x=x, w=w, d0=d0.T, d1=d1.T, d2=d2.T, gg1=g1.T, gg2=g2.T)

    def term_grid(self, t):
        nq = int(self.qoff[t + 1] - self.qoff[t])
        q = np.zeros((nq, self.D))
        lib().orc_sg_term_grid(self._h, t, _d(q))
        return q

    def grid_hh(self, lam=0.111803):
        V = np.zeros(self.sum_nq)
        lib().orc_sg_grid_hh(self._h, lam, _d(V))
        return V

    def hpsi(self, psi, V, ck=None, type_op=1, G=None, Jac=None, sqRJ=None, t0=0, t1=-1,
             tstride=1):
        """psi: (nb_ba, nvec) array.  Returns H.psi of the same shape."""
        psi = np.asarray(psi, dtype=np.float64)
        one = psi.ndim == 1
        if one:
            psi = psi[:, None]
        nvec = psi.shape[1]
        pf = np.asfortranarray(psi)
        out = np.zeros((self.nb_ba, nvec), order="F")
        ck = None if ck is None else np.ascontiguousarray(ck, dtype=np.float64)
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        V, G, Jac, sqRJ = c(V), c(G), c(Jac), c(sqRJ)
        rc = lib().orc_hpsi_range(self._h, type_op, _d(V), _d(ck), _d(G), _d(Jac), _d(sqRJ),
                                  nvec, _d(pf), _d(out), t0, t1, tstride)
        assert rc == 0
        return out[:, 0] if one else out


def _absbound(self, psi, V, ck=None):
    """sum_t |W_t r_t| per element (condition of the Smolyak scatter-sum)."""
    psi = np.asarray(psi, dtype=np.float64)
    one = psi.ndim == 1
    pf = np.asfortranarray(psi[:, None] if one else psi)
    out = np.zeros((self.nb_ba, pf.shape[1]), order="F")
    c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    V, ck = c(V), c(ck)
    assert lib().orc_hpsi_absbound(self._h, 1, _d(V), _d(ck), None, None, None, pf.shape[1],
                                   _d(pf), _d(out)) == 0
    return out[:, 0] if one else out


SG4.absbound = _absbound


def _hpsi_ld(self, psi, V, ck):
    """Type-1 H.psi in extended precision (long double contractions and sums)."""
    psi = np.asarray(psi, dtype=np.float64)
    one = psi.ndim == 1
    pf = np.asfortranarray(psi[:, None] if one else psi)
    out = np.zeros((self.nb_ba, pf.shape[1]), order="F")
    V = np.ascontiguousarray(V, dtype=np.float64)
    ck = np.ascontiguousarray(ck, dtype=np.float64)
    assert lib().orc_hpsi_ld(self._h, _d(V), _d(ck), pf.shape[1], _d(pf), _d(out)) == 0
    return out[:, 0] if one else out


SG4.hpsi_ld = _hpsi_ld


def _hpsi_t1g(self, psi, der, F):
    """General type-1 H.psi: sum_e F[e] D_e psi with der = [(i, j), ...]
    (0-based modes, -1 = none) and F of shape (nder, sum_nq)."""
    psi = np.asarray(psi, dtype=np.float64)
    one = psi.ndim == 1
    pf = np.asfortranarray(psi[:, None] if one else psi)
    out = np.zeros((self.nb_ba, pf.shape[1]), order="F")
    d = np.ascontiguousarray(np.asarray(der, dtype=np.int32).reshape(-1, 2))
    Fa = np.ascontiguousarray(F, dtype=np.float64).reshape(len(d), self.sum_nq)
    assert lib().orc_hpsi_t1g(self._h, len(d), d.ctypes.data_as(C.POINTER(C.c_int)), _d(Fa),
                              pf.shape[1], _d(pf), _d(out)) == 0
    return out[:, 0] if one else out


SG4.hpsi_t1g = _hpsi_t1g


def num_threads():
    return lib().orc_num_threads()


def henon_heiles(q, lam=0.111803):
    q = np.ascontiguousarray(q, dtype=np.float64)
    return lib().orc_henon_heiles(len(q), _d(q), lam)


def taylor_propagate(hfun, psi, dt, E0, npoly, poly_tol, eps=(0.0,), steps_per_segment=1,
                     per_state=False):
    """Taylor propagation (march_noD_field,
    SRC/sub_propagation/sub_module_propa_march.f90:1641-1842) with the field
    constant on each segment: hfun(X, eps) is the real H.psi of the columns of
    X under H0 + eps mu (complex states are applied as their real and
    imaginary parts, sub_OpPsi.f90:397-411).  Per step: w_0 = psi,
    w_{k+1} = (-i dt/(k+1)) (H - E0) w_k (:1756-1765), psi += w_{k+1} (:1787),
    cut after the first k whose ||w_{k+1}||^2 < poly_tol (:1791): for every
    state of the block together, or per state (per_state=True, the reference's
    loop over WP(j)); psi *= exp(-i E0 dt) (:1812).  psi: complex
    (nb_ba, ns).  Returns (psi(T), terms used per step and state)."""
    psi = np.array(psi, dtype=np.complex128)
    ns = psi.shape[1]
    used = []
    for e in eps:
        for _ in range(steps_per_segment):
            cols = [np.arange(ns)] if not per_state else [np.array([s]) for s in range(ns)]
            step_used = np.zeros(ns, dtype=int)
            for c in cols:
                w = psi[:, c].copy()
                for k in range(npoly):
                    hw = hfun(w.real, e) + 1j * hfun(w.imag, e)
                    w = (-1j * dt / (k + 1)) * (hw - E0 * w)
                    psi[:, c] += w
                    step_used[c] = k + 1
                    if np.max(np.sum(np.abs(w) ** 2, axis=0)) < poly_tol:
                        break
            psi *= np.exp(-1j * E0 * dt)
            used.append(step_used)
    return psi, np.array(used)
