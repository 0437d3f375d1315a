"""ctypes declarations of include/sg4hpsi.h."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsg4hpsi.so")

D_P = C.POINTER(C.c_double)
I_P = C.POINTER(C.c_int32)

PREC_FP64, PREC_F32, PREC_F32X2 = 0, 1, 2
OPT_GROUPING = 1
OPT_DETERMINISTIC = 2
# return codes (include/sg4hpsi.h)
OK, ERR_ARG, ERR_HIP, ERR_OOM, ERR_UNSUPPORTED, ERR_STATE = 0, -1, -2, -3, -4, -5
BASIS_HO, BASIS_LEGENDRE = 0, 1

# every symbol the header declares (checked by the CPU tests)
EXPORTS = [
    "sg4_last_error", "sg4_setup_create", "sg4_setup_sizes", "sg4_setup_tables",
    "sg4_setup_term_grid", "sg4_setup_basis1d", "sg4_setup_destroy", "sg4_setup_packed", "sg4_plan_create",
    "sg4_plan_create_batch",
    "sg4_plan_set_stream", "sg4_plan_set_operator", "sg4_plan_set_henon_heiles", "sg4_plan_set_jacobi_triatomic", "sg4_apply",
    "sg4_apply_host", "sg4_apply_rows", "sg4_plan_stats", "sg4_plan_timing",
    "sg4_plan_kernel_times", "sg4_grad_V", "sg4_plan_set_field", "sg4_plan_autotune", "sg4_plan_config", "sg4_plan_set_config",
    "sg4_apply_graph", "sg4_plan_set_option", "sg4_plan_split_info", "sg4_plan_mfma_stats",
    "sg4_plan_destroy", "sg4_apply_noalloc", "sg4_cheby_propagate", "sg4_cheby_field_gradient",
    "sg4_plan_grid_dipole", "sg4_plan_set_operator_type1", "sg4_taylor_propagate",
    "sg4_block_gram_work", "sg4_block_gram", "sg4_block_combine",
]


class SG4Error(RuntimeError):
    pass


class ModeDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nq_A", C.c_int32), ("nq_B", C.c_int32),
                ("nb_A", C.c_int32), ("nb_B", C.c_int32), ("Q0", C.c_double),
                ("scaleQ", C.c_double)]


class Tables(C.Structure):
    _fields_ = [("D", C.c_int32), ("Lmax", C.c_int32), ("nterms", C.c_int64),
                ("nb_ba", C.c_int64), ("levels", I_P), ("weights", D_P), ("map", I_P),
                ("nq", I_P), ("nb", I_P), ("mats", D_P), ("nodes", D_P), ("qweights", D_P)]


class ChebyParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("E0", C.c_double), ("Esc", C.c_double),
                ("ncoef", C.c_int32), ("coef", D_P), ("nseg", C.c_int32),
                ("steps_per_seg", C.c_int32), ("eps", D_P), ("ckpt_every", C.c_int32),
                ("prec", C.c_int32)]


class TaylorParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("E0", C.c_double), ("npoly", C.c_int32),
                ("poly_tol", C.c_double), ("nseg", C.c_int32), ("steps_per_seg", C.c_int32),
                ("eps", D_P), ("max_norm2", C.c_double), ("prec", C.c_int32)]


class Operator(C.Structure):
    _fields_ = [("type_op", C.c_int32), ("V", D_P), ("ck", D_P), ("G", D_P), ("Jac", D_P),
                ("sqRJ", D_P)]


_lib = None


def lib():
    """Load libsg4hpsi.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SG4Error("libsg4hpsi.so is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.sg4_last_error.restype = C.c_char_p
        L.sg4_setup_create.argtypes = [i32, i32, i32, C.POINTER(ModeDesc), C.POINTER(vp)]
        L.sg4_setup_sizes.argtypes = [vp] + [C.POINTER(i64)] * 4
        L.sg4_setup_tables.argtypes = [vp, C.POINTER(Tables)]
        L.sg4_setup_term_grid.argtypes = [vp, i64, D_P]
        L.sg4_setup_basis1d.argtypes = [vp, i32, i32, C.POINTER(i32), C.POINTER(i32)] + [D_P] * 5
        L.sg4_setup_destroy.argtypes = [vp]
        L.sg4_plan_create.argtypes = [C.POINTER(Tables), i32, C.POINTER(vp)]
        L.sg4_plan_create_batch.argtypes = [i32, C.POINTER(Tables), i32, C.POINTER(vp)]
        L.sg4_plan_set_stream.argtypes = [vp, vp]
        L.sg4_plan_set_operator.argtypes = [vp, C.POINTER(Operator)]
        L.sg4_plan_set_henon_heiles.argtypes = [vp, C.c_double]
        for f in ("sg4_apply", "sg4_apply_rows", "sg4_apply_noalloc"):
            getattr(L, f).argtypes = [vp, i32, vp, vp, i32]
        L.sg4_apply_host.argtypes = [vp, i32, D_P, D_P, i32]
        L.sg4_plan_stats.argtypes = [vp, i32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(i64)]
        L.sg4_plan_timing.argtypes = [vp, i32]
        L.sg4_plan_kernel_times.argtypes = [vp, C.POINTER(C.c_float), i32, C.POINTER(i32)]
        L.sg4_grad_V.argtypes = [vp, i32, vp, vp, vp]
        L.sg4_plan_set_field.argtypes = [vp, vp, C.c_double]
        L.sg4_plan_autotune.argtypes = [vp, i32, i32, C.POINTER(C.c_float)]
        L.sg4_plan_config.argtypes = [vp, i32, i32, C.POINTER(i32)]
        L.sg4_plan_set_config.argtypes = [vp, i32, i32, C.POINTER(i32)]
        L.sg4_apply_graph.argtypes = [vp, i32, vp, vp, i32]
        L.sg4_plan_set_option.argtypes = [vp, i32, i32]
        L.sg4_plan_set_jacobi_triatomic.argtypes = [vp, C.POINTER(C.c_double)]
        L.sg4_setup_packed.argtypes = [vp, C.POINTER(i32)]
        L.sg4_plan_split_info.argtypes = [vp, C.POINTER(i32), C.POINTER(C.c_int64)]
        L.sg4_plan_mfma_stats.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.sg4_plan_destroy.argtypes = [vp]
        L.sg4_plan_grid_dipole.argtypes = [vp, vp]
        L.sg4_plan_set_operator_type1.argtypes = [vp, i32, C.POINTER(i32), D_P]
        L.sg4_cheby_propagate.argtypes = [vp, vp, i32, vp, C.POINTER(ChebyParams)]
        L.sg4_cheby_field_gradient.argtypes = [vp, vp, vp, i32, vp, vp, i32,
                                               C.POINTER(C.c_int64), C.POINTER(ChebyParams),
                                               C.POINTER(C.c_double), D_P]
        L.sg4_taylor_propagate.argtypes = [vp, vp, i32, vp, C.POINTER(TaylorParams),
                                           C.POINTER(i32)]
        L.sg4_block_gram_work.argtypes = [i32, i32, C.POINTER(i64)]
        L.sg4_block_gram.argtypes = [vp, i32, i32, i64, vp, i64, vp, i64, vp, vp, i64]
        L.sg4_block_combine.argtypes = [vp, i32, i32, i64, vp, vp, i64, C.c_double, vp, i64]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().sg4_last_error()
        raise SG4Error(f"sg4 error {rc}: {msg.decode() if msg else ''}")
    return rc
