"""ctypes mirror of include/h2eri.h (the C-ABI of libh2eri_amd.so)."""
import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libh2eri_amd.so")

c_int_p = C.POINTER(C.c_int)
c_dbl_p = C.POINTER(C.c_double)


class Shell(C.Structure):
    _fields_ = [("am", C.c_int), ("nprim", C.c_int), ("x", C.c_double), ("y", C.c_double),
                ("z", C.c_double), ("alpha", c_dbl_p), ("coef", c_dbl_p)]


class DenseMat(C.Structure):
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("ld", C.c_int), ("size", C.c_int),
                ("data", c_dbl_p)]


class IntVec(C.Structure):
    _fields_ = [("capacity", C.c_int), ("length", C.c_int), ("data", c_int_p)]


DenseMatP = C.POINTER(DenseMat)
IntVecP = C.POINTER(IntVec)

_INT_FIELDS0 = ["natom", "nshell", "max_am", "max_shell_nbf", "num_bf", "num_sp", "num_sp_bfp",
                "pp_npts_layer", "pp_nlayer_ext"]
_IPTR_FIELDS0 = ["shell_bf_sidx", "sp_nbfp", "sp_bfp_sidx", "sp_shell_idx", "index_seq", "node_adm_pairs",
                 "node_adm_pairs_sidx", "node_inadm_pairs", "node_inadm_pairs_sidx", "plist", "plist_idx",
                 "plist_sidx", "dlist", "dlist_sidx"]
_INT_FIELDS1 = ["n_thread", "pt_dim", "xpt_dim", "krnl_dim", "QR_stop_type", "QR_stop_rank", "n_point",
                "krnl_mat_size", "max_leaf_points", "n_node", "root_idx", "n_leaf_node", "max_child",
                "max_neighbor", "max_level", "min_adm_level", "n_r_inadm_pair", "n_r_adm_pair", "n_UJ", "n_B",
                "n_D", "BD_JIT", "print_timers", "print_dbginfo"]
_IPTR_FIELDS1 = ["parent", "children", "pt_cluster", "mat_cluster", "n_child", "node_level", "node_height",
                 "level_n_node", "level_nodes", "height_n_node", "height_nodes", "r_inadm_pairs", "r_adm_pairs",
                 "node_inadm_lists", "node_n_r_inadm", "node_n_r_adm", "coord_idx", "B_p2i_rowptr",
                 "B_p2i_colidx", "B_p2i_val", "D_p2i_rowptr", "D_p2i_colidx", "D_p2i_val", "B_nrow", "B_ncol",
                 "D_nrow", "D_ncol"]


class H2ERIStruct(C.Structure):
    _fields_ = (
        [(n, C.c_int) for n in _INT_FIELDS0]
        + [(n, c_int_p) for n in _IPTR_FIELDS0]
        + [("thread_Kmat_workbuf", C.c_void_p), ("scr_tol", C.c_double), ("ext_tol", C.c_double),
           ("sp_center", c_dbl_p), ("sp_extent", c_dbl_p), ("box_extent", c_dbl_p), ("unc_denmat_x", c_dbl_p),
           ("H2_matvec_y", c_dbl_p), ("shells", C.POINTER(Shell)), ("sp_shells", C.POINTER(Shell)),
           ("sp", C.c_void_p), ("J_pair", C.POINTER(IntVecP)), ("J_row", C.POINTER(IntVecP)),
           ("ovlp_ff_idx", C.POINTER(IntVecP)), ("simint_buffs", C.c_void_p), ("eri_batch_buffs", C.c_void_p),
           ("c_B_blks", C.POINTER(DenseMatP)), ("c_D_blks", C.POINTER(DenseMatP)), ("thread_buffs", C.c_void_p)]
        + [(n, C.c_int) for n in _INT_FIELDS1]
        + [(n, c_int_p) for n in _IPTR_FIELDS1]
        + [("B_ptr", C.POINTER(C.c_size_t)), ("D_ptr", C.POINTER(C.c_size_t)), ("max_leaf_size", C.c_double),
           ("QR_stop_tol", C.c_double), ("coord", c_dbl_p), ("coord0", c_dbl_p), ("enbox", c_dbl_p),
           ("root_enbox", c_dbl_p), ("B_data", c_dbl_p), ("D_data", c_dbl_p), ("B_blk", C.c_void_p),
           ("D_blk0", C.c_void_p), ("D_blk1", C.c_void_p), ("U", C.POINTER(DenseMatP)),
           ("y0", C.POINTER(DenseMatP)), ("y1", C.POINTER(DenseMatP)), ("n_matvec", C.c_int),
           ("mat_size", C.c_size_t * 11), ("timers", C.c_double * 11), ("dev", C.c_void_p)]
    )


H2ERI_p = C.POINTER(H2ERIStruct)

# name -> (restype, argtypes); every function declared in include/h2eri.h
SIGNATURES = {
    "H2ERI_init": (None, [C.POINTER(H2ERI_p), C.c_double, C.c_double, C.c_double]),
    "H2ERI_destroy": (None, [H2ERI_p]),
    "H2ERI_print_statistic": (None, [H2ERI_p]),
    "H2E_reset_timers": (None, [H2ERI_p]),
    "CMS_read_mol_file": (None, [C.c_char_p, c_int_p, c_int_p, C.POINTER(C.POINTER(Shell))]),
    "CMS_destroy_shells": (None, [C.c_int, C.POINTER(Shell)]),
    "H2ERI_process_shells": (None, [H2ERI_p]),
    "H2ERI_partition": (None, [H2ERI_p]),
    "H2ERI_build_H2": (None, [H2ERI_p, C.c_int]),
    "H2ERI_matvec": (None, [H2ERI_p, c_dbl_p, c_dbl_p]),
    "H2ERI_build_Coulomb": (None, [H2ERI_p, c_dbl_p, c_dbl_p]),
    "H2ERI_build_exchange": (None, [H2ERI_p, c_dbl_p, c_dbl_p]),
    "H2ERI_exchange_workbuf_init": (None, [H2ERI_p]),
    "H2ERI_exchange_workbuf_free": (None, [H2ERI_p]),
    "H2ERI_build_exchange_dev": (None, [H2ERI_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "H2ERI_exchange_stats": (None, [H2ERI_p, c_dbl_p]),
    "H2ERI_exchange_plan_check": (C.c_int, [H2ERI_p, c_dbl_p]),
    "h2eri_last_error": (C.c_int, []),
    "h2eri_last_error_msg": (C.c_char_p, []),
    "h2eri_clear_error": (None, []),
    "H2ERI_set_precision": (None, [H2ERI_p, C.c_int]),
    "H2ERI_matvec_multi": (None, [H2ERI_p, C.c_int, c_dbl_p, c_dbl_p]),
    "H2ERI_build_Coulomb_multi": (None, [H2ERI_p, C.c_int, c_dbl_p, c_dbl_p]),
    "H2ERI_matvec_dev": (None, [H2ERI_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "H2ERI_matvec_multi_dev": (None, [H2ERI_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "H2ERI_build_Coulomb_dev": (None, [H2ERI_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "H2ERI_build_Coulomb_multi_dev": (None, [H2ERI_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "H2ERI_upload": (C.c_size_t, [H2ERI_p]),
    "H2ERI_graph_capture": (C.c_int, [H2ERI_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "H2ERI_graph_replay": (C.c_int, [H2ERI_p, C.c_void_p]),
    "H2ERI_save_rep": (C.c_int, [H2ERI_p, C.c_char_p]),
    "H2ERI_load_rep": (C.c_int, [H2ERI_p, C.c_char_p]),
    "H2ERI_build_H2_synthetic": (None, [H2ERI_p, C.c_ulonglong, C.c_int, C.c_double, C.c_double]),
    "H2ERI_partition_points_synthetic": (None, [H2ERI_p, C.c_int, c_dbl_p, c_int_p, C.c_int]),
    "H2ERI_set_phase_timing": (None, [H2ERI_p, C.c_int]),
    "H2ERI_get_phase_ms": (None, [H2ERI_p, c_dbl_p]),
    "H2ERI_get_launch_stats": (C.c_int, [H2ERI_p, C.c_int, c_dbl_p]),
    "H2ERI_matvec_bytes": (None, [H2ERI_p, c_dbl_p]),
    "H2ERI_set_tile_bytes": (None, [H2ERI_p, C.c_double]),
    "H2ERI_set_schedule": (None, [H2ERI_p, C.c_int]),
    "h2eri_struct_size": (C.c_size_t, []),
    "H2ERI_autotune": (C.c_int, [H2ERI_p, C.c_int, C.c_int, c_dbl_p]),
    "H2ERI_group_create": (C.c_void_p, [C.c_int, C.POINTER(H2ERI_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "H2ERI_group_matvec": (C.c_int, [C.c_void_p, C.c_void_p]),
    "H2ERI_group_destroy": (None, [C.c_void_p]),
    "H2ERI_group_status": (C.c_int, [C.c_void_p]),
    "H2ERI_group_set_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "H2ERI_group_autotune": (C.c_int, [C.c_void_p, C.c_int, c_dbl_p]),
    "H2ERI_plan_check": (C.c_int, [H2ERI_p, C.c_int]),
    "H2ERI_plan_emulate": (C.c_int, [H2ERI_p, C.c_int, C.c_int, c_dbl_p, c_dbl_p]),
    "H2ERI_set_executor": (None, [H2ERI_p, C.c_int]),
    "H2ERI_set_wave_tile_bytes": (None, [H2ERI_p, C.c_double]),
    "H2ERI_set_near_field_grid": (None, [H2ERI_p, C.c_int]),
    "H2ERI_set_exchange_scratch": (None, [H2ERI_p, C.c_double]),
    "H2ERI_set_block_compression": (None, [H2ERI_p, C.c_int]),
    "H2ERI_jit_emulate": (C.c_int, [H2ERI_p, c_dbl_p, c_dbl_p]),
    "H2ERI_jit_stats": (C.c_int, [H2ERI_p, c_dbl_p]),
    "H2ERI_get_executor": (C.c_int, [H2ERI_p]),
    "H2ERI_device_status": (C.c_int, [H2ERI_p]),
    "H2ERI_set_spin_limit": (None, [H2ERI_p, C.c_int]),
    "H2ERI_plan_stats": (C.c_int, [H2ERI_p, C.c_int, C.c_int, c_dbl_p]),
    "H2ERI_exchange_dlist": (None, [H2ERI_p, c_dbl_p]),
    "H2ERI_matvec_dev_noalloc": (None, [H2ERI_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "H2ERI_direct_matvec_rows": (None, [H2ERI_p, C.c_int, C.POINTER(C.c_int), c_dbl_p, c_dbl_p]),
    "H2ERI_queue_stats": (C.c_int, [H2ERI_p, C.c_int, c_dbl_p]),
    "H2ERI_queue_emulate": (C.c_int, [H2ERI_p, C.c_int, C.c_int, C.c_int, c_dbl_p, c_dbl_p]),
}

_lib = None


def load():
    """Loads libh2eri_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    # One HIP runtime per process: when PyTorch is present its libamdhip64
    # must be loaded first, so that the library's libamdhip64.so.7 dependency
    # resolves to the same runtime (PyTorch's own libraries ask for
    # "libamdhip64.so", which would otherwise map a second runtime).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if lib.h2eri_struct_size() != C.sizeof(H2ERIStruct):
        raise RuntimeError("struct H2ERI layout mismatch between library and binding")
    _lib = lib
    return lib
