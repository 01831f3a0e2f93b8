"""ctypes binding of libffno_hip.so (include/ffno_hip.h). Loading fails loudly: there is no fallback path."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libffno_hip.so")

c_i64 = ctypes.c_int64
c_ptr = ctypes.c_void_p


class FfnoRows(ctypes.Structure):
    _fields_ = [("B", ctypes.c_int), ("voxels", c_i64), ("C", ctypes.c_int), ("nwg", ctypes.c_int)]


class FfnoStatsOut(ctypes.Structure):
    _fields_ = [("depth_w", c_ptr), ("partials", c_ptr)]


class FfnoAssemble(ctypes.Structure):
    _fields_ = [("base", c_ptr), ("t", c_ptr), ("aff_a", c_ptr), ("aff_b", c_ptr), ("scale", c_ptr),
                ("act_gelu", ctypes.c_int)]


class FfnoAssembleT(ctypes.Structure):
    _fields_ = [("base_dot", c_ptr), ("t", c_ptr), ("t_dot", c_ptr), ("aff_a", c_ptr), ("aff_b", c_ptr),
                ("tan_b", c_ptr), ("tan_c", c_ptr), ("scale", c_ptr), ("act_gelu", ctypes.c_int)]


class FfnoTanStatsOut(ctypes.Structure):
    _fields_ = [("y", c_ptr), ("stats", c_ptr), ("depth_w", c_ptr), ("partials", c_ptr)]


class FfnoHeadDesc(ctypes.Structure):
    _fields_ = [("n_species", ctypes.c_int), ("levels", ctypes.c_int * 4), ("log10_mass", ctypes.c_float * 4),
                ("log_g", ctypes.c_float * 32), ("ev", ctypes.c_float * 32), ("ndxi", ctypes.c_float * 32),
                ("dion", ctypes.c_float * 32), ("mean_x", ctypes.c_float * 6), ("std_x", ctypes.c_float * 6),
                ("mean_y", ctypes.c_float * 32), ("std_y", ctypes.c_float * 32), ("t_channel", ctypes.c_int),
                ("ne_channel", ctypes.c_int), ("rho_channel", ctypes.c_int), ("total_bound", ctypes.c_float),
                ("ln_ccon", ctypes.c_float), ("ln_k", ctypes.c_float), ("e_over_k", ctypes.c_float),
                ("dxi_coef", ctypes.c_float)]


class FfnoTwGemm(ctypes.Structure):
    _fields_ = [("a_hi", c_ptr), ("a_lo", c_ptr), ("m", ctypes.c_int), ("k", ctypes.c_int), ("n", ctypes.c_int),
                ("m_pad", ctypes.c_int), ("k_pad", ctypes.c_int), ("batch", ctypes.c_int), ("x", c_ptr),
                ("y", c_ptr), ("b_split", c_i64), ("x_b_hi", c_i64), ("x_b_lo", c_i64), ("y_b_hi", c_i64),
                ("y_b_lo", c_i64), ("k_split", c_i64), ("x_k_hi", c_i64), ("x_k_lo", c_i64), ("m_split", c_i64),
                ("y_m_hi", c_i64), ("y_m_lo", c_i64), ("accumulate", ctypes.c_int), ("a_l3", c_ptr),
                ("twiddle", c_ptr), ("tw_n1", ctypes.c_int), ("tw_n2", ctypes.c_int), ("tw_cols", c_i64),
                ("tw_bsplit", c_i64)]


class FfnoColGemm(ctypes.Structure):
    _fields_ = [("a_hi", c_ptr), ("a_lo", c_ptr), ("a_l3", c_ptr), ("m", ctypes.c_int), ("k", ctypes.c_int),
                ("m_pad", ctypes.c_int), ("k_pad", ctypes.c_int), ("B", ctypes.c_int), ("hw", c_i64),
                ("C", ctypes.c_int), ("x", c_ptr), ("y", c_ptr)]


class FfnoMix(ctypes.Structure):
    _fields_ = [("re", c_ptr), ("im", c_ptr), ("C", ctypes.c_int), ("rows", ctypes.c_int), ("modes", ctypes.c_int),
                ("rank", ctypes.c_int), ("mode_stride", c_i64), ("r_split", c_i64), ("r_hi", c_i64),
                ("r_lo", c_i64), ("wc", c_ptr), ("gate", c_ptr), ("v_p", c_ptr), ("u_p", c_ptr),
                ("factor", c_ptr), ("parts", ctypes.c_int)]


class FfnoZGrad(ctypes.Structure):
    _fields_ = [("B", ctypes.c_int), ("hw", c_i64), ("D", ctypes.c_int), ("R", ctypes.c_int), ("C", ctypes.c_int),
                ("r_pad", ctypes.c_int), ("d_pad", ctypes.c_int), ("xv", c_ptr), ("xbar_spec", c_ptr),
                ("yv_bar", c_ptr), ("y_spec", c_ptr), ("tables", c_ptr), ("zbar", c_ptr)]


class FfnoGnBack(ctypes.Structure):
    _fields_ = [("t", c_ptr), ("stats", c_ptr), ("gamma", c_ptr), ("aff_a", c_ptr), ("aff_b", c_ptr),
                ("scale", c_ptr), ("act_gelu", ctypes.c_int), ("sums", c_ptr), ("depth_w", c_ptr), ("wbar", c_ptr)]


# entry point -> argument types (every one returns int)
SIGNATURES = {
    "ffno_lift": [FfnoRows, c_ptr, c_ptr, c_ptr, FfnoStatsOut, c_ptr],
    "ffno_gn_finalize": [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr],
    "ffno_block_in": [FfnoRows, FfnoAssemble, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr],
    "ffno_post": [FfnoRows, c_ptr, c_ptr, c_ptr, c_ptr, FfnoStatsOut, c_ptr],
    "ffno_post_gn_noalloc": [FfnoRows, c_ptr, c_ptr, c_ptr, c_ptr] + [c_ptr] * 6 + [c_ptr],
    "ffno_fuse": [FfnoRows, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr,
                  FfnoStatsOut, c_ptr, c_ptr],
    "ffno_resmlp": [FfnoRows, FfnoAssemble, c_ptr, c_ptr, c_ptr, c_ptr, FfnoStatsOut, c_ptr],
    "ffno_head": [FfnoRows, FfnoAssemble, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, ctypes.POINTER(FfnoHeadDesc), c_ptr,
                  c_ptr],
    "ffno_tw_gemm": [ctypes.POINTER(FfnoTwGemm), c_ptr],
    "ffno_mix": [ctypes.POINTER(FfnoMix), c_ptr],
    "ffno_col_gemm": [ctypes.POINTER(FfnoColGemm), c_ptr],
    "ffno_gn_grad_stats": [FfnoRows, FfnoGnBack, c_ptr, c_ptr, c_ptr],
    "ffno_gn_grad_finalize": [ctypes.c_int, ctypes.c_int, c_ptr, c_ptr, c_ptr],
    "ffno_mlp_bwd": [FfnoRows, c_ptr, FfnoGnBack, FfnoAssemble, c_ptr, c_ptr, c_ptr, ctypes.c_int, c_ptr, c_ptr,
                     c_ptr, c_ptr],
    "ffno_fuse_bwd": [FfnoRows, c_ptr, FfnoGnBack] + [c_ptr] * 18 + [c_ptr],
    "ffno_fuse_mlp_bwd": [FfnoRows, c_ptr, FfnoGnBack] + [c_ptr] * 12 + [c_ptr],
    "ffno_fuse_gates_bwd": [FfnoRows] + [c_ptr] * 14 + [c_ptr, c_ptr, c_ptr, c_ptr] + [c_ptr],
    "ffno_gates_bwd": [FfnoRows] + [c_ptr] * 15 + [c_ptr],
    "ffno_head_bwd": [FfnoRows] + [c_ptr] * 7 + [ctypes.POINTER(FfnoHeadDesc)] + [c_ptr] * 3 + [c_ptr],
    "ffno_lift_bwd": [FfnoRows, c_ptr, FfnoGnBack, c_ptr, c_ptr, c_ptr],
    "ffno_dot_reduce": [FfnoRows, ctypes.c_int, c_ptr, ctypes.c_int, c_ptr, c_ptr, c_ptr],
    "ffno_zgrad_columns": [ctypes.POINTER(FfnoZGrad), c_ptr],
    "ffno_gn_tangent_stats": [FfnoRows] + [c_ptr] * 5 + [c_ptr],
    "ffno_gn_tangent_finalize": [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [c_ptr] * 5 + [c_ptr],
    "ffno_assemble_t": [FfnoRows, FfnoAssembleT, c_ptr, c_ptr],
    "ffno_lift_t": [FfnoRows, c_ptr, c_ptr, c_ptr, c_ptr],
    "ffno_gates_t": [FfnoRows] + [c_ptr] * 8 + [c_ptr],
    "ffno_mlp_t": [FfnoRows, FfnoAssemble, c_ptr, c_ptr, c_ptr, c_ptr, FfnoTanStatsOut, c_ptr],
    "ffno_fuse_t": [FfnoRows, c_ptr, c_ptr, FfnoAssembleT, FfnoAssembleT] + [c_ptr] * 6 + [FfnoTanStatsOut, c_ptr],
    "ffno_head_t": [FfnoRows, FfnoAssemble] + [c_ptr] * 6 + [ctypes.POINTER(FfnoHeadDesc), c_ptr, c_ptr],
    "ffno_lowrank_tables": [c_i64, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [c_ptr] * 5 +
                           [ctypes.c_int, ctypes.c_int, c_ptr, c_ptr, c_ptr],
    "ffno_head_output": [ctypes.c_int, c_i64, c_ptr, c_ptr, ctypes.POINTER(FfnoHeadDesc), c_ptr, c_ptr],
    "ffno_head_output_bwd": [ctypes.c_int, c_i64, c_ptr, c_ptr, ctypes.POINTER(FfnoHeadDesc), c_ptr, c_ptr, c_ptr,
                             c_ptr],
}

_lib = None


class FfnoError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it has not been built (python -c 'import __graft_entry__ as g; g.build()')."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FfnoError(f"{LIB_PATH} is missing: build the HIP library first (__graft_entry__.build())")
        handle = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        handle.ffno_last_error.restype = ctypes.c_char_p
        handle.ffno_version.restype = ctypes.c_int
        _lib = handle
    return _lib


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise FfnoError(f"{name} failed ({rc}): {lib().ffno_last_error().decode()}")
