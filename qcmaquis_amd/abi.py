"""ctypes mirror of include/qcmaquis_amd.h and the loader of the in-tree library.

The product path runs only through the HIP library: if ``libqcmaquis_amd.so`` is missing the
import of :func:`lib` raises; there is no CPU fallback.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libqcmaquis_amd.so")

MQ_OK = 0
MQ_ERR_INVALID = -1
MQ_ERR_HIP = -2
MQ_ERR_NOMEM = -3
MQ_ERR_UNSUPPORTED = -4
MQ_SYM_TWOU1 = 1
MQ_SYM_SU2U1 = 2
MQ_PREC_FP64 = 0
MQ_PREC_FP32 = 1
MQ_PREC_BF16X3 = 2
MQ_PREC_FP64_PLAIN = 3
MQ_NOALLOC_SCRATCH_ELEMS = 1 << 23
MQ_OPT_PLAIN_TABLES = 1
MQ_OPT_SMALL_TILE_MAX = 2
MQ_SMALL_TILE_AUTO = -1
MQ_OPT_CHUNK_MB = 3
MQ_CHUNK_AUTO = -1
MQ_OPT_TERM_SOURCES = 4
MQ_OPT_SBLOCK_PANEL = 5
MQ_SBLOCK_PANEL_AUTO = -1
MQ_OPT_SBLOCK_XCD = 6
MQ_OPT_AUTOTUNE = 7
MQ_OPT_JCD_GMRES = 8
MQ_OPT_OVERLAP = 9
MQ_QC_GRADIENT = 1
MQ_RIGHT_PAIRED = 0
MQ_LEFT_PAIRED = 1
MQ_NO_PARTNER = 0xFFFFFFFF

P = C.POINTER


class IndexDesc(C.Structure):
    _fields_ = [("n", C.c_int32), ("charges", P(C.c_int32)), ("sizes", P(C.c_uint32))]


class MPODesc(C.Structure):
    _fields_ = [
        ("n_ops", C.c_int32),
        ("op_spin", P(C.c_int32)),
        ("op_block_ptr", P(C.c_int64)),
        ("block_charges", P(C.c_int32)),
        ("block_entry_ptr", P(C.c_int64)),
        ("entry_rc", P(C.c_uint32)),
        ("entry_spins", P(C.c_int32)),
        ("entry_coeff", P(C.c_double)),
        ("row_dim", C.c_uint32),
        ("col_dim", C.c_uint32),
        ("n_terms", C.c_int64),
        ("term_b", P(C.c_uint32)),
        ("term_op", P(C.c_uint32)),
        ("term_scale", P(C.c_double)),
        ("left_spin", P(C.c_int32)),
        ("right_spin", P(C.c_int32)),
        ("left_herm", P(C.c_uint32)),
        ("left_phase", P(C.c_int32)),
        ("right_herm", P(C.c_uint32)),
        ("right_phase", P(C.c_int32)),
    ]


class BIndexDesc(C.Structure):
    _fields_ = [
        ("bra", IndexDesc),
        ("ket", IndexDesc),
        ("aux_dim", C.c_uint32),
        ("n_cohorts", C.c_int32),
        ("cohort_lr", P(C.c_uint32)),
        ("offsets", P(C.c_int64)),
        ("conj_scales", P(C.c_double)),
        ("transposes", P(C.c_uint8)),
        ("tr", P(C.c_uint8)),
    ]


class BoundaryInfo(C.Structure):
    _fields_ = [("n_cohorts", C.c_int32), ("aux_dim", C.c_uint32), ("n_bra", C.c_int32), ("n_ket", C.c_int32),
                ("total_elements", C.c_uint64)]


class HpsiStats(C.Structure):
    _fields_ = [
        ("psi_elements", C.c_uint64),
        ("n_t_tasks", C.c_uint64),
        ("n_cohorts", C.c_uint64),
        ("n_lines", C.c_uint64),
        ("n_terms", C.c_uint64),
        ("flops_total", C.c_uint64),
        ("flops_gemm", C.c_uint64),
        ("flops_gemm_exec", C.c_uint64),
        ("bytes_compulsory", C.c_uint64),
        ("t_elements", C.c_uint64),
        ("s_elements", C.c_uint64),
        ("l_panel_elements", C.c_uint64),
        ("r_tr_elements", C.c_uint64),
        ("build_seconds", C.c_double),
        ("s_line_elements", C.c_uint64),
        ("gathered_elements", C.c_uint64),
        ("n_sblocks", C.c_uint64),
        ("union_elements", C.c_uint64),
        ("seconds_tables", C.c_double),
        ("seconds_blocks", C.c_double),
        ("seconds_plain", C.c_double),
        ("small_lines", C.c_uint64),
        ("small_terms", C.c_uint64),
        ("small_gathered", C.c_uint64),
        ("max_union_tiles", C.c_uint64),
        ("split_reduces", C.c_uint64),
        ("sblock_macs", C.c_uint64),
        ("n_chunks", C.c_uint64),
        ("ts_buffer_elements", C.c_uint64),
        ("gemm_policy_t", C.c_int32),
        ("gemm_policy_c", C.c_int32),
        ("seconds_tune", C.c_double),
    ]


class DavidsonStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("residual_norm", C.c_double), ("seconds", C.c_double),
                ("hpsi_seconds", C.c_double), ("refine_iterations", C.c_int32)]


class GrowthStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("n_lines", "n_terms", "n_sblocks", "t_elements", "s_elements",
                                          "out_elements", "flops_total", "flops_gemm")] + \
               [(f, C.c_double) for f in ("seconds_tables", "seconds_blocks", "seconds_upload", "seconds_device",
                                          "seconds_total")]


vp = C.c_void_p
i32 = C.c_int
dp = P(C.c_double)

# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ("mq_version", i32, []),
    ("mq_abi_struct_sizes", i32, [P(C.c_uint32), i32]),
    ("mq_ctx_create", i32, [i32, i32, P(vp)]),
    ("mq_ctx_destroy", i32, [vp]),
    ("mq_last_error", C.c_char_p, [vp]),
    ("mq_ctx_sync", i32, [vp]),
    ("mq_ctx_stream", vp, [vp]),
    ("mq_ctx_set_option", i32, [vp, i32, i32]),
    ("mq_mpo_create", i32, [vp, P(MPODesc), P(vp)]),
    ("mq_mpo_destroy", i32, [vp]),
    ("mq_boundary_create_initial", i32, [vp, P(IndexDesc), C.c_uint32, P(vp)]),
    ("mq_boundary_upload", i32, [vp, P(BIndexDesc), dp, P(vp)]),
    ("mq_boundary_info", i32, [vp, P(BoundaryInfo)]),
    ("mq_boundary_export_index", i32, [vp, P(C.c_int32), P(C.c_uint32), P(C.c_int32), P(C.c_uint32),
                                       P(C.c_uint32), P(C.c_int64), dp, P(C.c_uint8), P(C.c_uint8)]),
    ("mq_boundary_download", i32, [vp, dp]),
    ("mq_boundary_destroy", i32, [vp]),
    ("mq_boundary_growth_stats", i32, [vp, P(GrowthStats)]),
    ("mq_mps_upload", i32, [vp, P(IndexDesc), P(IndexDesc), P(IndexDesc), i32, dp, P(vp)]),
    ("mq_mps_elements", i32, [vp, P(C.c_uint64)]),
    ("mq_mps_download", i32, [vp, i32, dp]),
    ("mq_mps_destroy", i32, [vp]),
    ("mq_lstep", i32, [vp, vp, vp, vp, vp, i32, i32, P(vp)]),
    ("mq_rstep", i32, [vp, vp, vp, vp, vp, i32, i32, P(vp)]),
    ("mq_hpsi_plan_create", i32, [vp, vp, vp, vp, P(IndexDesc), P(IndexDesc), P(IndexDesc), P(vp)]),
    ("mq_hpsi_plan_create_batch", i32, [vp, i32, P(vp), P(vp), P(vp), P(IndexDesc), P(IndexDesc), P(IndexDesc),
                                         P(vp)]),
    ("mq_hpsi_plan_stats", i32, [vp, P(HpsiStats)]),
    ("mq_hpsi_plan_psi_elements", i32, [vp, P(C.c_uint64)]),
    ("mq_hpsi", i32, [vp, vp, vp, i32]),
    ("mq_hpsi_timed", i32, [vp, vp, vp, i32, P(C.c_float)]),
    ("mq_hpsi_host", i32, [vp, dp, dp, i32]),
    ("mq_hpsi_plan_use_graphs", i32, [vp, i32]),
    ("mq_hpsi_term_gradient", i32, [vp, dp, dp, C.c_int64]),
    ("mq_ctx_tuned", i32, [vp, i32, P(C.c_uint64), P(i32), P(i32), P(C.c_float), P(C.c_float)]),
    ("mq_hpsi_plan_destroy", i32, [vp]),
    ("mq_davidson", i32, [vp, dp, i32, P(dp), C.c_double, i32, i32, dp, dp, P(DavidsonStats)]),
    ("mq_davidson_noalloc", i32, [vp, dp, i32, P(dp), C.c_double, i32, i32, dp, dp, P(DavidsonStats)]),
    ("mq_davidson_batch", i32, [vp, dp, C.c_double, i32, i32, dp, dp, P(C.c_int32), P(DavidsonStats)]),
    ("mq_block_svd", i32, [vp, i32, P(C.c_int32), P(C.c_int32), dp, dp, dp, dp]),
    ("mq_boundary_gram", i32, [vp, i32, C.c_double, i32, P(C.c_int32), P(C.c_uint32), dp]),
    ("mq_qc_build", i32, [i32, i32, dp, dp, C.c_double, C.c_double, P(vp)]),
    ("mq_qc_build_ex", i32, [i32, i32, dp, dp, C.c_double, C.c_double, i32, P(vp)]),
    ("mq_qc_integral_gradient", i32, [vp, P(dp), dp, dp]),
    ("mq_qc_info", i32, [vp, P(i32), P(C.c_uint32), dp]),
    ("mq_qc_site_mpo", i32, [vp, i32, i32, P(MPODesc)]),
    ("mq_qc_phys", i32, [vp, i32, P(IndexDesc)]),
    ("mq_qc_destroy", i32, [vp]),
    ("mq_wigner_6j", C.c_double, [i32] * 6),
    ("mq_wigner_9j", C.c_double, [i32] * 9),
    ("mq_index_position", i32, [P(IndexDesc), C.c_int32, C.c_int32]),
]

_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libqcmaquis_amd.so is not built; run `python -m qcmaquis_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        # the library writes these structs: a layout mismatch would overwrite Python memory
        mine = [C.sizeof(t) for t in (IndexDesc, MPODesc, BIndexDesc, BoundaryInfo, HpsiStats, DavidsonStats,
                                       GrowthStats)]
        theirs = (C.c_uint32 * len(mine))()
        got = L.mq_abi_struct_sizes(theirs, len(mine))
        if got != len(mine) or list(theirs) != mine:
            raise ImportError("libqcmaquis_amd.so struct layouts %s do not match the Python binding %s; rebuild"
                              % (list(theirs)[:got], mine))
        _lib = L
    return _lib
