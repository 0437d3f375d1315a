"""elvibrot_amd -- MI355X (gfx950) Smolyak SG4 H.psi for ElVibRot-style drivers.

Host-side mirror of the reference's operator interface for the SG4 path:

* :class:`SG4Basis` -- the SG4 basis set-up (RecSparseGrid_ForDP_type4,
  SRC/sub_Basis/sub_quadra_SparseBasis.f90:922-1434): 1-D bases, term set,
  Smolyak weights, packed basis and map.
* :class:`SG4Operator` -- the device plan of one operator; its
  :meth:`SG4Operator.apply` is sub_TabOpPsi_FOR_SGtype4(Psi, OpPsi, para_Op)
  (SRC/sub_Operator/sub_OpPsi_SG4.f90:681).
* :func:`sub_TabOpPsi_FOR_SGtype4` -- the same call with the reference's
  argument order.

Everything goes through the C ABI of ``libsg4hpsi.so`` (include/sg4hpsi.h).
There is no CPU fallback: if the HIP library is missing or the device is not
available, the calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import _capi
from ._capi import SG4Error, lib, PREC_FP64, PREC_F32, PREC_F32X2, BASIS_HO, BASIS_LEGENDRE

__all__ = ["SG4Basis", "SG4Operator", "SG4Batch", "SG4Error", "sub_TabOpPsi_FOR_SGtype4", "lib",
           "PREC_FP64", "PREC_F32", "PREC_F32X2", "BASIS_HO", "BASIS_LEGENDRE",
           "henon_heiles_basis"]


class SG4Basis:
    """Smolyak SG type 4 basis with ``D`` modes, grid level ``L`` and basis level ``LB``.

    ``modes`` is a list of dicts with keys kind, nq_A, nq_B, nb_A, nb_B, Q0,
    scaleQ (the &basis_nD fields); the defaults are the HO basis with
    nq(l) = nb(l) = 1 + 2 l used by the Henon-Heiles tests of the reference
    (Working_tests/MPI_tests/6D_Davidson_S0/shell_run:128-134).
    """

    def __init__(self, D, L, LB=None, modes=None):
        LB = L if LB is None else LB
        if modes is None:
            modes = [dict()] * D
        arr = (_capi.ModeDesc * D)()
        for k, m in enumerate(modes):
            arr[k].kind = m.get("kind", BASIS_HO)
            arr[k].nq_A = m.get("nq_A", 1)
            arr[k].nq_B = m.get("nq_B", 2)
            arr[k].nb_A = m.get("nb_A", arr[k].nq_A)
            arr[k].nb_B = m.get("nb_B", arr[k].nq_B)
            arr[k].Q0 = m.get("Q0", 0.0)
            arr[k].scaleQ = m.get("scaleQ", 1.0)
        self.D, self.L, self.LB = D, L, LB
        self._modes = arr
        h = C.c_void_p()
        _capi.check(lib().sg4_setup_create(D, L, LB, arr, C.byref(h)))
        self._h = h
        n = [C.c_int64() for _ in range(4)]
        _capi.check(lib().sg4_setup_sizes(h, *[C.byref(x) for x in n]))
        self.nterms, self.nb_ba, self.sum_nb, self.sum_nq = (x.value for x in n)
        self.tables = _capi.Tables()
        _capi.check(lib().sg4_setup_tables(h, C.byref(self.tables)))

    def __del__(self):
        # at interpreter exit the module globals may already be gone; the
        # process is ending then and the library frees nothing that matters
        if getattr(self, "_h", None) and lib is not None:
            try:
                lib().sg4_setup_destroy(self._h)
            except Exception:
                pass
            self._h = None

    # numpy copies of the tables (host memory owned by the set-up)
    def levels(self):
        return np.ctypeslib.as_array(self.tables.levels, (self.nterms, self.D)).copy()

    def weights(self):
        return np.ctypeslib.as_array(self.tables.weights, (self.nterms,)).copy()

    def map(self):
        return np.ctypeslib.as_array(self.tables.map, (self.sum_nb,)).copy()

    def packed(self):
        """(nb_ba, D) int32: 0-based mode indices of every packed row."""
        out = np.zeros((self.nb_ba, self.D), dtype=np.int32)
        _capi.check(lib().sg4_setup_packed(self._h, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def term_grid(self, t):
        lv = np.ctypeslib.as_array(self.tables.levels, (self.nterms, self.D))[t]
        nq = 1
        for k in range(self.D):
            nq *= self.basis1d(k, int(lv[k]))["nq"]
        q = np.zeros((nq, self.D))
        _capi.check(lib().sg4_setup_term_grid(self._h, t, q.ctypes.data_as(_capi.D_P)))
        return q

    def basis1d(self, k, l):
        nq, nb = C.c_int32(), C.c_int32()
        _capi.check(lib().sg4_setup_basis1d(self._h, k, l, C.byref(nq), C.byref(nb),
                                            None, None, None, None, None))
        nq, nb = nq.value, nb.value
        x, w = np.zeros(nq), np.zeros(nq)
        d = [np.zeros((nb, nq)) for _ in range(3)]
        p = lambda a: a.ctypes.data_as(_capi.D_P)
        _capi.check(lib().sg4_setup_basis1d(self._h, k, l, C.byref(C.c_int32()),
                                            C.byref(C.c_int32()), p(x), p(w), p(d[0]), p(d[1]),
                                            p(d[2])))
        return dict(nq=nq, nb=nb, x=x, w=w, d0=d[0].T, d1=d[1].T, d2=d[2].T)


def henon_heiles_basis(D, L):
    """SG4 HO basis of the reference Henon-Heiles tests: nq(l)=nb(l)=1+2l, Q0=0, scaleQ=1."""
    return SG4Basis(D, L, L)


class SG4Operator:
    """Device plan of an SG4 operator (``para_Op`` for the SG4 path).

    potential:
      * ``("henon_heiles", lam)`` -- V evaluated on the device grids with
        kinetic -1/2 sum d^2 (type 1);
      * ``("jacobi_triatomic", (mu_R, mu_r, D, a, Re, re, c))`` -- the type-10
        triatomic model, V, G, Jac evaluated on the device grids;
      * a numpy array ``V`` of length ``basis.sum_nq`` with ``ck`` (type 1) or
        with ``G``, ``Jac``, ``sqRJ`` (type 10);
      * ``("type1", der, F)`` -- the general type-1 operator
        sum_e F[e](q) D_e with derivative terms der[e] = (i, j), -1 = none.
    """

    def __init__(self, basis, max_nvec=16, potential=("henon_heiles", 0.111803), ck=None,
                 stream=None, type_op=1, G=None, Jac=None, sqRJ=None, der=None, F=None):
        self.basis = basis
        self.max_nvec = max_nvec
        self.type_op = type_op
        h = C.c_void_p()
        _capi.check(lib().sg4_plan_create(C.byref(basis.tables), max_nvec, C.byref(h)))
        self._h = h
        if stream is not None:
            _capi.check(lib().sg4_plan_set_stream(h, C.c_void_p(stream)))
        if isinstance(potential, tuple) and potential[0] == "henon_heiles":
            _capi.check(lib().sg4_plan_set_henon_heiles(h, float(potential[1])))
            return
        if isinstance(potential, tuple) and potential[0] == "type1":
            # general type 1: der = [(i, j), ...] (-1 = none), F: (nder, sum_nq) grids
            d = np.ascontiguousarray(np.asarray(potential[1], dtype=np.int32).reshape(-1, 2))
            Fa = np.ascontiguousarray(potential[2], dtype=np.float64).reshape(len(d), basis.sum_nq)
            _capi.check(lib().sg4_plan_set_operator_type1(
                h, len(d), d.ctypes.data_as(C.POINTER(C.c_int32)), Fa.ctypes.data_as(_capi.D_P)))
            self.type_op = 11
            return
        if isinstance(potential, tuple) and potential[0] == "jacobi_triatomic":
            prm = (C.c_double * 7)(*[float(x) for x in potential[1]])
            _capi.check(lib().sg4_plan_set_jacobi_triatomic(h, prm))
            self.type_op = 10
            return
        c64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        V = c64(potential)
        if V.shape != (basis.sum_nq,):
            raise ValueError("V must have basis.sum_nq entries")
        op = _capi.Operator()
        op.type_op = type_op
        op.V = V.ctypes.data_as(_capi.D_P)
        keep = [V]
        if type_op == 1:
            ckv = c64(-0.5 * np.ones(basis.D) if ck is None else ck)
            op.ck = ckv.ctypes.data_as(_capi.D_P)
            keep.append(ckv)
        else:
            G, Jac, sqRJ = c64(G), c64(Jac), c64(sqRJ)
            if G.size != basis.sum_nq * basis.D ** 2 or Jac.size != basis.sum_nq or \
                    sqRJ.size != basis.sum_nq:
                raise ValueError("G, Jac, sqRJ must be given on every grid point")
            op.G = G.ctypes.data_as(_capi.D_P)
            op.Jac = Jac.ctypes.data_as(_capi.D_P)
            op.sqRJ = sqRJ.ctypes.data_as(_capi.D_P)
            keep += [G, Jac, sqRJ]
        _capi.check(lib().sg4_plan_set_operator(h, C.byref(op)))

    def __del__(self):
        if getattr(self, "_h", None) and lib is not None:
            try:
                lib().sg4_plan_destroy(self._h)
            except Exception:
                pass
            self._h = None

    @property
    def nb_ba(self):
        return self.basis.nb_ba

    def apply_host(self, psi, prec=PREC_FP64):
        """H.psi for a host block psi (nb_ba,) or (nb_ba, nvec); returns numpy."""
        psi = np.asarray(psi, dtype=np.float64)
        one = psi.ndim == 1
        P = np.asfortranarray(psi[:, None] if one else psi)
        nvec = P.shape[1]
        out = np.zeros_like(P, order="F")
        _capi.check(lib().sg4_apply_host(self._h, nvec, P.ctypes.data_as(_capi.D_P),
                                         out.ctypes.data_as(_capi.D_P), prec))
        return out[:, 0] if one else out

    def apply(self, psi, out, prec=PREC_FP64):
        """H.psi on device memory.  ``psi``/``out``: torch tensors of shape (nvec, nb_ba)
        (row-major), i.e. the column-major nb_ba x nvec block of the reference."""
        nvec = psi.shape[0]
        _capi.check(lib().sg4_apply(self._h, nvec, C.c_void_p(psi.data_ptr()),
                                    C.c_void_p(out.data_ptr()), prec))
        return out

    def apply_rows(self, psi, out, prec=PREC_FP64):
        """H.psi in the interleaved device layout: tensors of shape (nb_ba, nvec)."""
        nvec = psi.shape[1]
        _capi.check(lib().sg4_apply_rows(self._h, nvec, C.c_void_p(psi.data_ptr()),
                                         C.c_void_p(out.data_ptr()), prec))
        return out

    def grad_V(self, phi, psi, out):
        """d<phi_v|H|psi_v>/dV(p) for every grid point p (device tensors in the
        rows layout: phi, psi (nb_ba, nvec); out (sum_nq, nvec))."""
        nvec = psi.shape[1]
        _capi.check(lib().sg4_grad_V(self._h, nvec, C.c_void_p(phi.data_ptr()),
                                     C.c_void_p(psi.data_ptr()), C.c_void_p(out.data_ptr())))
        return out

    def set_field(self, mu, eps):
        """V = V0 + eps * mu (mu: device tensor with sum_nq entries, or None)."""
        _capi.check(lib().sg4_plan_set_field(self._h, None if mu is None else
                                             C.c_void_p(mu.data_ptr()), float(eps)))

    def autotune(self, nvec, prec=PREC_FP64):
        """Pick the fastest launch parameters for blocks of nvec vectors; returns ms."""
        ms = C.c_float()
        _capi.check(lib().sg4_plan_autotune(self._h, nvec, prec, C.byref(ms)))
        return ms.value

    def config(self, nvec, prec=PREC_FP64):
        c = (C.c_int32 * 4)()
        _capi.check(lib().sg4_plan_config(self._h, nvec, prec, c))
        return dict(nthreads=c[0], lds_budget=c[1], grid=c[2], xcd=c[3])

    def set_config(self, nvec, prec=PREC_FP64, nthreads=256, lds_budget=65536, grid=0, xcd=0):
        """Launch parameters for (nvec, prec), e.g. from an earlier autotune."""
        c = (C.c_int32 * 4)(nthreads, lds_budget, grid, xcd)
        _capi.check(lib().sg4_plan_set_config(self._h, nvec, prec, c))

    def apply_graph(self, psi, out, prec=PREC_FP64):
        """apply_rows through a captured HIP graph (rows layout tensors)."""
        nvec = psi.shape[1]
        _capi.check(lib().sg4_apply_graph(self._h, nvec, C.c_void_p(psi.data_ptr()),
                                          C.c_void_p(out.data_ptr()), prec))
        return out

    def set_grouping(self, on=True):
        """Register-blocked mode groups and, where it applies, the split path
        (kinetic part as one sparse matrix; default), or one pass per mode with
        the kinetic part on the grid (the reference's structure)."""
        _capi.check(lib().sg4_plan_set_option(self._h, _capi.OPT_GROUPING, 1 if on else 0))

    def set_deterministic(self, on=True):
        """Order-independent (bitwise reproducible) sum over the terms: 128-bit
        fixed-point accumulation on the split type-1 path."""
        _capi.check(lib().sg4_plan_set_option(self._h, _capi.OPT_DETERMINISTIC, 1 if on else 0))

    def split_info(self):
        """(split path in use, stored entries of the kinetic matrix)"""
        sp, nnz = C.c_int32(), C.c_int64()
        _capi.check(lib().sg4_plan_split_info(self._h, C.byref(sp), C.byref(nnz)))
        return bool(sp.value), nnz.value

    def timing(self, enable=True):
        _capi.check(lib().sg4_plan_timing(self._h, 1 if enable else 0))

    def kernel_times(self, max_n=100000):
        buf = (C.c_float * max_n)()
        n = C.c_int32()
        _capi.check(lib().sg4_plan_kernel_times(self._h, buf, max_n, C.byref(n)))
        return np.array(buf[: n.value], dtype=np.float64)

    def grid_dipole(self):
        """mu = sum_k q_k on every grid point (device tensor of sum_nq values)."""
        import torch
        mu = torch.empty(self.basis.sum_nq, dtype=torch.float64, device="cuda")
        _capi.check(lib().sg4_plan_grid_dipole(self._h, C.c_void_p(mu.data_ptr())))
        return mu

    def mfma_stats(self):
        """(flops the matrix cores executed in the last split-precision apply's
        term launch, number of group matrices; -1 before the first such apply)"""
        f, n = C.c_double(), C.c_int64()
        _capi.check(lib().sg4_plan_mfma_stats(self._h, C.byref(f), C.byref(n)))
        return f.value, n.value

    def stats(self, nvec):
        f, b, n = C.c_double(), C.c_double(), C.c_int64()
        _capi.check(lib().sg4_plan_stats(self._h, nvec, C.byref(f), C.byref(b), C.byref(n)))
        return dict(flops=f.value, bytes=b.value, items=n.value)


class SG4Batch(SG4Operator):
    """Several independent SG4 operators in one plan (block-diagonal over the
    concatenated packed rows; one launch per apply for all instances).

    potential: ("henon_heiles", lam) for all instances, or a list of V grids
    (one per instance) with a list of ck arrays."""

    def __init__(self, bases, max_nvec=16, potential=("henon_heiles", 0.111803), ck=None,
                 stream=None):
        self.bases = list(bases)
        self.max_nvec = max_nvec
        self.type_op = 1
        n = len(self.bases)
        arr = (_capi.Tables * n)()
        for i, b in enumerate(self.bases):
            arr[i] = b.tables
        h = C.c_void_p()
        _capi.check(lib().sg4_plan_create_batch(n, arr, max_nvec, C.byref(h)))
        self._h = h
        self.row0 = np.cumsum([0] + [b.nb_ba for b in self.bases])
        self.q0 = np.cumsum([0] + [b.sum_nq for b in self.bases])
        self.basis = _BatchShape(self.bases, self.row0, self.q0)
        if stream is not None:
            _capi.check(lib().sg4_plan_set_stream(h, C.c_void_p(stream)))
        if isinstance(potential, tuple) and potential[0] == "henon_heiles":
            _capi.check(lib().sg4_plan_set_henon_heiles(h, float(potential[1])))
            return
        V = np.ascontiguousarray(np.concatenate(potential), dtype=np.float64)
        ckv = np.ascontiguousarray(np.concatenate(
            [(-0.5 * np.ones(b.D)) for b in self.bases] if ck is None else ck), dtype=np.float64)
        op = _capi.Operator()
        op.type_op = 1
        op.V = V.ctypes.data_as(_capi.D_P)
        op.ck = ckv.ctypes.data_as(_capi.D_P)
        _capi.check(lib().sg4_plan_set_operator(h, C.byref(op)))


class _BatchShape:
    """What the drivers read from ``op.basis`` for a batch."""

    def __init__(self, bases, row0, q0):
        self.bases = bases
        self.nb_ba = int(row0[-1])
        self.sum_nq = int(q0[-1])
        self.D = sum(b.D for b in bases)
        self.nterms = sum(b.nterms for b in bases)

    def term_grid(self, t):
        for b in self.bases:
            if t < b.nterms:
                return b.term_grid(t)
            t -= b.nterms
        raise IndexError(t)


def sub_TabOpPsi_FOR_SGtype4(Psi, OpPsi, para_Op, prec=PREC_FP64):
    """OpPsi = H.Psi for a block (reference argument order,
    SRC/sub_Operator/sub_OpPsi_SG4.f90:681).  Psi/OpPsi are torch device
    tensors (nvec, nb_ba) or numpy arrays (nb_ba, nvec)."""
    if isinstance(Psi, np.ndarray):
        OpPsi[...] = para_Op.apply_host(Psi, prec)
        return OpPsi
    return para_Op.apply(Psi, OpPsi, prec)
