"""Python host layer over libcqmm.so (ctypes).

It mirrors the Fortran host API the library replaces:
  * ``System`` is the single-process stand-in for CONQUEST's index set-up
    (make_cs / matrix_ini / immi, src/cover_module.f90:133-429,
    src/matrix_elements_module.f90:67-142, src/mult_module.f90:271-1102).
  * ``Plan`` + ``Plan.mat_mult`` is ``mat_mult(myid,a,lena,b,lenb,c,lenc,a_b_c)``
    (src/multiply_module.f90:115) for one process, type 1 or 2.

The library is loaded from ``conquest-release_amd/lib/libcqmm.so``; a missing
library raises instead of falling back to any CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libcqmm.so")

CQMM_DEVICE_PTRS = 1
CQMM_FP32 = 2
CQMM_FP32_ACC64 = 8
CQMM_FP32_RUN64 = 16
CQMM_BF16X3 = 32
CQMM_NO_SYNC = 4

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)


class Halo(C.Structure):
    _fields_ = [
        ("np_in_halo", C.c_int32), ("ni_in_halo", C.c_int32),
        ("nh_part", i32p), ("j_beg", i32p), ("lab_hcell", i32p), ("lab_hcover", i32p),
        ("j_seq", i32p), ("ndimj", i32p), ("i_h2d", i32p), ("i_halo", i32p), ("i_hbeg", i32p),
        ("ndimi", i32p), ("len_i_halo", C.c_int64), ("len_i_hbeg", C.c_int64),
    ]


class Trans(C.Structure):
    _fields_ = [("n_hnab", i32p), ("i_beg", i32p), ("i_nd_beg", i32p), ("i_prim", i32p),
                ("len_i_prim", C.c_int64)]


class BIndex(C.Structure):
    _fields_ = [("n_parts", C.c_int32), ("array_posn", i32p), ("n_atoms", i32p),
                ("part_nabs", i32p), ("nd_offset", i32p), ("bindex", i32p),
                ("len_bindex", C.c_int64)]


class MultDesc(C.Structure):
    _fields_ = [
        ("mult_type", C.c_int32), ("n_prim", C.c_int32), ("ncovery", C.c_int32), ("ncoverz", C.c_int32),
        ("ahalo", Halo), ("chalo", Halo), ("ltrans", Trans), ("bmat", BIndex),
        ("i_cc2seq", i32p), ("n_cell_parts", C.c_int32),
        ("lena", C.c_int64), ("lenb", C.c_int64), ("lenc", C.c_int64),
    ]


class PlanInfo(C.Structure):
    _fields_ = [("n_triplets", C.c_int64), ("n_out_blocks", C.c_int64), ("n_out_total", C.c_int64),
                ("flops", C.c_double), ("uniform_nsf", C.c_int32), ("mult_type", C.c_int32),
                ("kernel", C.c_int32), ("n_tiles", C.c_int32), ("n_entries", C.c_int64), ("n_loads", C.c_int64),
                ("tile_family", C.c_int32), ("tile_wgs", C.c_int32), ("n_mirror", C.c_int64)]


KERNEL_AUTO, KERNEL_TRIPLET, KERNEL_TILE = 0, 1, 2
TILES_DENSE4, TILES_DENSE_MIXED, TILES_MIXED, TILES_CLASS, TILES_UNIFORM = 1, 2, 3, 5, 16


class RangeInfo(C.Structure):
    _fields_ = [("rcut", C.c_double), ("length", C.c_int64), ("n_blocks", C.c_int64),
                ("ni_in_halo", C.c_int32), ("np_in_halo", C.c_int32), ("max_nab", C.c_int32),
                ("fits_int32", C.c_int32)]


class MatrixIndex(C.Structure):
    _fields_ = [("n_parts", C.c_int32), ("n_atoms", i32p), ("nab_offset", i32p), ("nd_offset", i64p),
                ("n_prim", C.c_int32), ("n_nab", i32p), ("i_acc", i32p), ("i_nd_acc", i32p), ("ndimi", i32p),
                ("onsite", i64p), ("i_part", i32p), ("i_seq", i32p), ("ndimj", i32p), ("n_nabs", C.c_int64),
                ("length", C.c_int64)]


class McwSetup(C.Structure):
    _fields_ = [("range_L", C.c_int32), ("range_S", C.c_int32), ("range_H", C.c_int32), ("range_LS", C.c_int32),
                ("range_LSL", C.c_int32), ("mult_L_S_LS", C.c_int32), ("mult_LS_L_LSL", C.c_int32),
                ("mult_LSL_SL_L", C.c_int32), ("max_iter", C.c_int32), ("precision", C.c_int32),
                ("n_final_fp64", C.c_int32), ("fp32_tol", C.c_double)]


class McwResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("fp32_iterations", C.c_int32), ("stop_reason", C.c_int32),
                ("restored", C.c_int32), ("energy", C.c_double), ("electrons", C.c_double), ("c", C.c_double),
                ("seconds", C.c_double), ("energy_trace", C.c_double * 256)]


class LnvSetup(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "range_L", "range_S", "range_H", "range_LS", "range_LH", "range_LSL", "range_SLS",
        "mult_L_S_LS", "mult_L_H_LH", "mult_HL_S_LSL", "mult_HLS_LS_L", "mult_SL_S_SLS", "mult_SLS_LS_L",
        "mult_LS_L_LSL", "mult_LSL_SL_H")]


class DmmSetup(C.Structure):
    _fields_ = [("lnv", LnvSetup), ("range_T", C.c_int32), ("range_TL", C.c_int32), ("mult_T_L_TL", C.c_int32),
                ("mult_TL_T_L", C.c_int32), ("n_L_iterations", C.c_int32), ("vary_mu", C.c_int32),
                ("maxpulay", C.c_int32), ("global_tolerance", C.c_int32), ("n_atoms", C.c_int32),
                ("tolerance", C.c_double), ("min_pulay_step", C.c_double), ("max_pulay_step", C.c_double),
                ("lin_tol", C.c_double), ("electrons_target", C.c_double), ("delta_e", C.c_double)]


class DmmResult(C.Structure):
    _fields_ = [("done", C.c_int32), ("early_iterations", C.c_int32), ("late_iterations", C.c_int32),
                ("inflexions", C.c_int32), ("energy", C.c_double), ("electrons", C.c_double),
                ("residual", C.c_double), ("seconds", C.c_double)]


LNV_K, LNV_E, LNV_M3, LNV_PHI = 1, 2, 4, 8


class InitMcwSetup(C.Structure):
    _fields_ = [("range_T", C.c_int32), ("range_S", C.c_int32), ("range_H", C.c_int32), ("range_L", C.c_int32),
                ("range_TS", C.c_int32), ("range_TH", C.c_int32), ("mult_T_S_TS", C.c_int32),
                ("mult_T_H_TH", C.c_int32), ("mult_TH_T_L", C.c_int32), ("n_electrons", C.c_double)]


class HotellingSetup(C.Structure):
    _fields_ = [("range_T", C.c_int32), ("range_S", C.c_int32), ("range_TS", C.c_int32),
                ("mult_T_S_TS", C.c_int32), ("mult_TS_T_T", C.c_int32), ("max_iter", C.c_int32),
                ("tol", C.c_double)]


class SysInfo(C.Structure):
    _fields_ = [("natoms", C.c_int32), ("n_parts", C.c_int32), ("ng_cover", C.c_int32),
                ("mx_mcover", C.c_int32), ("ncover", C.c_int32 * 3), ("nspanl", C.c_int32 * 3),
                ("ngrid", C.c_int32 * 3), ("box", C.c_double * 3)]


class InfoGlobal(C.Structure):
    _fields_ = [("natoms", C.c_int32), ("numprocs", C.c_int32), ("npcell", C.c_int32 * 3),
                ("rcell", C.c_double * 3), ("md_step", C.c_int32), ("has_velocity", C.c_int32),
                ("has_md_step", C.c_int32), ("index", C.c_int32)]


# Symbols declared in include/cqmm.h and include/cqmm_system.h.
EXPORTS = [
    "cqmm_last_error", "cqmm_init", "cqmm_finalize", "cqmm_device_alloc", "cqmm_device_free",
    "cqmm_memcpy_h2d", "cqmm_memcpy_d2h", "cqmm_device_copy", "cqmm_synchronize", "cqmm_timer_start", "cqmm_timer_stop",
    "cqmm_plan_create", "cqmm_plan_create_symmetric", "cqmm_plan_destroy", "cqmm_mat_mult",
    "cqmm_mat_mult_fp32_noalloc", "cqmm_sys_plan_create_symmetric",
    "cqmm_noalloc_scratch_bytes", "cqmm_plan_get_info",
    "cqmm_plan_get_triplets", "cqmm_plan_set_kernel", "cqmm_plan_set_tile_workgroups", "cqmm_plan_tune",
    "cqmm_sys_create", "cqmm_sys_destroy", "cqmm_sys_add_range", "cqmm_sys_range_info",
    "cqmm_sys_get_info", "cqmm_sys_mult_desc", "cqmm_sys_blocks", "cqmm_sys_primary_order",
    "cqmm_sys_fill", "cqmm_random_structure", "cqmm_sys_plan_create", "cqmm_sys_matrix_index",
    "cqmm_sys_transpose_pairs", "cqmm_matrix2_file_name", "cqmm_sys_dump_matrix2", "cqmm_sys_load_matrix2",
    "cqmm_info_global_file_name", "cqmm_sys_dump_info_global", "cqmm_load_info_global",
    "cqmm_matrix_file_name", "cqmm_dump_matrix_text", "cqmm_grab_matrix_text",
    "cqmm_ctx_create", "cqmm_ctx_destroy", "cqmm_ctx_add_range", "cqmm_ctx_set_transpose", "cqmm_ctx_add_mult",
    "cqmm_ctx_set_product_precision", "cqmm_ctx_get_product_precision",
    "cqmm_matrix_allocate", "cqmm_matrix_free", "cqmm_matrix_length", "cqmm_matrix_range",
    "cqmm_matrix_device_ptr", "cqmm_matrix_set", "cqmm_matrix_get", "cqmm_matrix_product", "cqmm_matrix_sum",
    "cqmm_matrix_scale", "cqmm_matrix_product_trace", "cqmm_matrix_trace", "cqmm_matrix_symmetrise",
    "cqmm_matrix_transpose", "cqmm_mcweeny", "cqmm_mcweeny_batch", "cqmm_init_mcweeny", "cqmm_hotelling",
    "cqmm_mcweeny_spin", "cqmm_init_mcweeny_spin",
    "cqmm_lnv", "cqmm_correct_electrons", "cqmm_dmm",
    "cqmm_purify_default_params", "cqmm_purify_hbm_estimate", "cqmm_purify", "cqmm_serve",
]

_lib = None


class CqmmError(RuntimeError):
    pass


def lib():
    """Load libcqmm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CqmmError(f"libcqmm.so not found at {LIB_PATH}; run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    L.cqmm_last_error.restype = C.c_char_p
    L.cqmm_random_structure.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_uint64, f64p, f64p]
    L.cqmm_sys_create.argtypes = [C.c_int32, f64p, f64p, i32p, i32p, C.c_double, C.POINTER(C.c_void_p)]
    L.cqmm_sys_destroy.argtypes = [C.c_void_p]
    L.cqmm_sys_add_range.argtypes = [C.c_void_p, C.c_double, i32p]
    L.cqmm_sys_range_info.argtypes = [C.c_void_p, C.c_int32, C.POINTER(RangeInfo)]
    L.cqmm_sys_get_info.argtypes = [C.c_void_p, C.POINTER(SysInfo)]
    L.cqmm_sys_mult_desc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(MultDesc)]
    L.cqmm_sys_blocks.argtypes = [C.c_void_p, C.c_int32, i32p, i32p, i32p, i32p, i64p, i32p, i32p]
    L.cqmm_sys_primary_order.argtypes = [C.c_void_p, i32p]
    L.cqmm_sys_fill.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_uint64, C.c_uint64, f64p, C.c_int32, f64p]
    L.cqmm_matrix2_file_name.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_char_p, C.c_int32]
    L.cqmm_sys_dump_matrix2.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(f64p), C.c_char_p, C.c_int32]
    L.cqmm_sys_load_matrix2.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32, C.c_int32, C.POINTER(f64p),
                                        C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.cqmm_info_global_file_name.argtypes = [C.c_int32, C.c_char_p, C.c_int32]
    L.cqmm_sys_dump_info_global.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int32, f64p]
    L.cqmm_load_info_global.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(InfoGlobal), i32p, f64p, f64p]
    L.cqmm_matrix_file_name.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_char_p, C.c_int32]
    L.cqmm_dump_matrix_text.argtypes = [C.c_char_p, f64p, C.c_int64]
    L.cqmm_grab_matrix_text.argtypes = [C.c_char_p, f64p, C.c_int64]
    L.cqmm_init.argtypes = [C.c_int]
    L.cqmm_plan_create.argtypes = [C.POINTER(MultDesc), C.POINTER(C.c_void_p)]
    L.cqmm_plan_create_symmetric.argtypes = [C.POINTER(MultDesc), C.c_int64, C.POINTER(C.c_int64),
                                             C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                             C.POINTER(C.c_void_p)]
    L.cqmm_plan_destroy.argtypes = [C.c_void_p]
    L.cqmm_plan_get_info.argtypes = [C.c_void_p, C.POINTER(PlanInfo)]
    L.cqmm_plan_get_triplets.argtypes = [C.c_void_p, i64p, i64p, i64p]
    L.cqmm_plan_set_kernel.argtypes = [C.c_void_p, C.c_int32]
    L.cqmm_plan_set_tile_workgroups.argtypes = [C.c_void_p, C.c_int32]
    L.cqmm_plan_tune.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                 C.c_uint32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.cqmm_mat_mult.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                C.c_int64, C.c_uint32]
    L.cqmm_mat_mult_fp32_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                C.c_int64, C.c_uint32]
    L.cqmm_noalloc_scratch_bytes.restype = C.c_int64
    L.cqmm_device_alloc.argtypes = [C.POINTER(C.c_void_p), C.c_int64]
    L.cqmm_device_free.argtypes = [C.c_void_p]
    L.cqmm_device_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    L.cqmm_memcpy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    L.cqmm_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    L.cqmm_timer_stop.argtypes = [C.POINTER(C.c_float)]
    vp = C.c_void_p
    L.cqmm_sys_plan_create.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(vp)]
    L.cqmm_sys_plan_create_symmetric.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(vp)]
    L.cqmm_sys_matrix_index.argtypes = [vp, C.c_int32, C.POINTER(MatrixIndex)]
    L.cqmm_sys_transpose_pairs.argtypes = [vp, C.c_int32, i64p, i64p, i64p, i32p, i32p]
    L.cqmm_ctx_create.argtypes = [C.POINTER(vp)]
    L.cqmm_ctx_destroy.argtypes = [vp]
    L.cqmm_ctx_add_range.argtypes = [vp, C.c_int32, C.POINTER(MatrixIndex)]
    L.cqmm_ctx_set_transpose.argtypes = [vp, C.c_int32, C.c_int64, i64p, i64p, i32p, i32p]
    L.cqmm_ctx_add_mult.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp]
    L.cqmm_ctx_set_product_precision.argtypes = [vp, C.c_int32]
    L.cqmm_ctx_get_product_precision.argtypes = [vp, i32p]
    L.cqmm_matrix_allocate.argtypes = [vp, C.c_int32, i32p]
    L.cqmm_matrix_free.argtypes = [vp, C.c_int32]
    L.cqmm_matrix_length.argtypes = [vp, C.c_int32, i64p]
    L.cqmm_matrix_range.argtypes = [vp, C.c_int32, i32p]
    L.cqmm_matrix_device_ptr.argtypes = [vp, C.c_int32, C.POINTER(vp)]
    L.cqmm_matrix_set.argtypes = [vp, C.c_int32, vp]
    L.cqmm_matrix_get.argtypes = [vp, C.c_int32, vp]
    L.cqmm_matrix_product.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.cqmm_matrix_sum.argtypes = [vp, C.c_double, C.c_int32, C.c_double, C.c_int32]
    L.cqmm_matrix_scale.argtypes = [vp, C.c_double, C.c_int32]
    L.cqmm_matrix_product_trace.argtypes = [vp, C.c_int32, C.c_int32, f64p]
    L.cqmm_matrix_trace.argtypes = [vp, C.c_int32, f64p]
    L.cqmm_matrix_symmetrise.argtypes = [vp, C.c_int32]
    L.cqmm_mcweeny.argtypes = [vp, C.POINTER(McwSetup), C.c_int32, C.c_int32, C.c_int32, C.POINTER(McwResult)]
    L.cqmm_mcweeny_batch.argtypes = [C.c_int32, C.POINTER(vp), C.POINTER(McwSetup), i32p, i32p, i32p, C.c_uint32,
                                     C.POINTER(McwResult), i32p]
    L.cqmm_init_mcweeny.argtypes = [vp, C.POINTER(InitMcwSetup), C.c_int32, C.c_int32, C.c_int32, C.c_int32, f64p]
    L.cqmm_hotelling.argtypes = [vp, C.POINTER(HotellingSetup), C.c_int32, C.c_int32, i32p, f64p]
    L.cqmm_mcweeny_spin.argtypes = [vp, C.POINTER(McwSetup), C.c_int32, i32p, i32p, i32p, C.POINTER(McwResult)]
    L.cqmm_init_mcweeny_spin.argtypes = [vp, C.POINTER(InitMcwSetup), C.c_int32, i32p, i32p, i32p, i32p, f64p,
                                         f64p]
    _lib = L
    return L


def check(status, what=""):
    if status != 0:
        msg = lib().cqmm_last_error().decode(errors="replace")
        raise CqmmError(f"{what}: status {status}: {msg}")


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def matrix2_file_name(root, index=0, inode=1):
    """dump_matrix2's file name: <root>matrix2.i<index>.p<inode>."""
    buf = C.create_string_buffer(4096)
    check(lib().cqmm_matrix2_file_name(str(root).encode(), index, inode, buf, 4096), "cqmm_matrix2_file_name")
    return buf.value.decode()


def info_global_file_name(index=0):
    """dump_InfoMatGlobal's file name: InfoGlobal.i<index>."""
    buf = C.create_string_buffer(256)
    check(lib().cqmm_info_global_file_name(index, buf, 256), "cqmm_info_global_file_name")
    return buf.value.decode()


def load_info_global(path, index=0, natoms_max=1 << 24):
    """Read an InfoGlobal file: dict of its scalars plus glob_to_node, coords
    (natoms, 3) and velocity (natoms, 3, or None when the file has none)."""
    head = InfoGlobal()
    check(lib().cqmm_load_info_global(str(path).encode(), index, 0, C.byref(head), None, None, None),
          "cqmm_load_info_global")
    natoms = head.natoms
    if natoms > natoms_max:
        raise CqmmError(f"load_info_global: {path} holds {natoms} atoms, room for {natoms_max}")
    nodes = np.zeros(max(natoms, 1), np.int32)
    xyz = np.zeros((max(natoms, 1), 3), np.float64)
    vel = np.zeros((max(natoms, 1), 3), np.float64)
    check(lib().cqmm_load_info_global(str(path).encode(), index, natoms, C.byref(head), _p(nodes, C.c_int32),
                                      _p(xyz, C.c_double), _p(vel, C.c_double)), "cqmm_load_info_global")
    n = head.natoms
    return dict(natoms=n, numprocs=head.numprocs, npcell=tuple(head.npcell), rcell=tuple(head.rcell),
                md_step=head.md_step, has_md_step=bool(head.has_md_step), index=head.index,
                glob_to_node=nodes[:n].copy(), coords=xyz[:n].copy(),
                velocity=vel[:n].copy() if head.has_velocity else None)


def matrix_file_name(stub, numprocs=1, inode=1):
    """dump_matrix's file name: <stub>matrix.<inode, zero-padded>."""
    buf = C.create_string_buffer(256)
    check(lib().cqmm_matrix_file_name(str(stub).encode(), numprocs, inode, buf, 256), "cqmm_matrix_file_name")
    return buf.value.decode()


def dump_matrix_text(path, values):
    """dump_matrix: one f25.18 value per line, in storage order."""
    v = np.ascontiguousarray(values, np.float64)
    check(lib().cqmm_dump_matrix_text(str(path).encode(), _p(v, C.c_double), v.size), "cqmm_dump_matrix_text")


def grab_matrix_text(path, length):
    """grab_matrix: `length` values from a dump_matrix file."""
    v = np.zeros(length, np.float64)
    check(lib().cqmm_grab_matrix_text(str(path).encode(), _p(v, C.c_double), length), "cqmm_grab_matrix_text")
    return v


def random_structure(natoms, density=0.05, dmin=2.0, seed=1):
    """Seeded uniform positions in a periodic cube (Angstrom); returns (xyz, edge)."""
    xyz = np.zeros((natoms, 3), dtype=np.float64)
    edge = C.c_double(0.0)
    check(lib().cqmm_random_structure(natoms, density, dmin, seed, _p(xyz, C.c_double), C.byref(edge)),
          "cqmm_random_structure")
    return xyz, edge.value


def default_grid(edge, target=5.0):
    n = max(1, int(round(edge / target)))
    return (n, n, n)


class System:
    """Single-process CONQUEST index set built from atom positions."""

    def __init__(self, xyz, box, nsf, rcut_cover, grid=None):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.natoms = xyz.shape[0]
        box = np.ascontiguousarray(np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)))
        nsf = np.ascontiguousarray(np.broadcast_to(np.asarray(nsf, dtype=np.int32), (self.natoms,)))
        if grid is None:
            grid = default_grid(float(box.min()))
        g = np.ascontiguousarray(np.asarray(grid, dtype=np.int32))
        self.xyz, self.box, self.nsf, self.grid = xyz, box, nsf, g
        self.rcut_cover = float(rcut_cover)
        h = C.c_void_p()
        check(lib().cqmm_sys_create(self.natoms, _p(xyz, C.c_double), _p(box, C.c_double), _p(nsf, C.c_int32),
                                    _p(g, C.c_int32), self.rcut_cover, C.byref(h)), "cqmm_sys_create")
        self.h = h
        self.ranges = []

    def close(self):
        if self.h:
            lib().cqmm_sys_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_range(self, rcut):
        rid = C.c_int32()
        check(lib().cqmm_sys_add_range(self.h, float(rcut), C.byref(rid)), "cqmm_sys_add_range")
        self.ranges.append(float(rcut))
        return rid.value

    def range_info(self, rid):
        info = RangeInfo()
        check(lib().cqmm_sys_range_info(self.h, rid, C.byref(info)), "cqmm_sys_range_info")
        return info

    def info(self):
        info = SysInfo()
        check(lib().cqmm_sys_get_info(self.h, C.byref(info)), "cqmm_sys_get_info")
        return info

    def length(self, rid):
        return self.range_info(rid).length

    def transpose_pairs(self, rid):
        """Transpose pairs of range rid: (p, q, n1, n2), p <= q 1-based block
        offsets of blocks (i,j,s) and (j,i,-s), n1 x n2 the dimensions of p."""
        n = C.c_int64()
        check(lib().cqmm_sys_transpose_pairs(self.h, rid, C.byref(n), None, None, None, None),
              "cqmm_sys_transpose_pairs")
        p, q = np.zeros(n.value, np.int64), np.zeros(n.value, np.int64)
        a, b = np.zeros(n.value, np.int32), np.zeros(n.value, np.int32)
        check(lib().cqmm_sys_transpose_pairs(self.h, rid, C.byref(n), _p(p, C.c_int64), _p(q, C.c_int64),
                                             _p(a, C.c_int32), _p(b, C.c_int32)), "cqmm_sys_transpose_pairs")
        return p, q, a, b

    def mult_desc(self, mult_type, ra, rb, rc):
        """Descriptor of a product; its arrays point into this system's memory,
        so the descriptor keeps the system alive."""
        d = MultDesc()
        check(lib().cqmm_sys_mult_desc(self.h, mult_type, ra, rb, rc, C.byref(d)), "cqmm_sys_mult_desc")
        d._system = self
        return d

    def blocks(self, rid):
        """Block list in storage order as a dict of numpy arrays."""
        n = self.range_info(rid).n_blocks
        out = dict(prim=np.zeros(n, np.int32), gid_i=np.zeros(n, np.int32), gid_j=np.zeros(n, np.int32),
                   shift=np.zeros((n, 3), np.int32), offset=np.zeros(n, np.int64),
                   nd1=np.zeros(n, np.int32), nd2=np.zeros(n, np.int32))
        check(lib().cqmm_sys_blocks(self.h, rid, _p(out["prim"], C.c_int32), _p(out["gid_i"], C.c_int32),
                                    _p(out["gid_j"], C.c_int32), _p(out["shift"], C.c_int32),
                                    _p(out["offset"], C.c_int64), _p(out["nd1"], C.c_int32),
                                    _p(out["nd2"], C.c_int32)), "cqmm_sys_blocks")
        return out

    def primary_order(self):
        out = np.zeros(self.natoms, np.int32)
        check(lib().cqmm_sys_primary_order(self.h, _p(out, C.c_int32)), "cqmm_sys_primary_order")
        return out

    def dump_matrix2(self, rid, values, path, binary=True):
        """Write range `rid` arrays (one per spin) in CONQUEST's dump_matrix2 format."""
        arrs = [np.ascontiguousarray(v, np.float64) for v in values]
        for a in arrs:
            if a.size != self.length(rid):
                raise CqmmError("dump_matrix2: length mismatch")
        ptrs = (C.POINTER(C.c_double) * len(arrs))(*[_p(a, C.c_double) for a in arrs])
        check(lib().cqmm_sys_dump_matrix2(self.h, rid, len(arrs), ptrs, str(path).encode(), int(bool(binary))),
              "cqmm_sys_dump_matrix2")

    def load_matrix2(self, rid, path, binary=True, nspin_max=2):
        """Read a dump_matrix2 file into range `rid`: (list of arrays, blocks loaded, blocks skipped)."""
        arrs = [np.zeros(self.length(rid), np.float64) for _ in range(nspin_max)]
        ptrs = (C.POINTER(C.c_double) * nspin_max)(*[_p(a, C.c_double) for a in arrs])
        ns, nl, nk = C.c_int32(), C.c_int64(), C.c_int64()
        check(lib().cqmm_sys_load_matrix2(self.h, rid, str(path).encode(), int(bool(binary)), nspin_max, ptrs,
                                          C.byref(ns), C.byref(nl), C.byref(nk)), "cqmm_sys_load_matrix2")
        return arrs[:ns.value], nl.value, nk.value

    def dump_info_global(self, path, index=0, md_step=None, velocity=None):
        """Write InfoGlobal (dump_InfoMatGlobal) for this system; velocity (natoms, 3) by global id."""
        vp = None
        if velocity is not None:
            vel = np.ascontiguousarray(velocity, np.float64)
            if vel.shape != (self.natoms, 3):
                raise CqmmError("dump_info_global: velocity must be (natoms, 3)")
            vp = _p(vel, C.c_double)
        check(lib().cqmm_sys_dump_info_global(self.h, str(path).encode(), index, int(md_step is not None),
                                              int(md_step or 0), vp), "cqmm_sys_dump_info_global")

    def fill(self, rid, kind=0, seed=1, tag=0, params=(2.0,)):
        v = np.zeros(self.length(rid), np.float64)
        p = np.ascontiguousarray(np.asarray(params, np.float64))
        check(lib().cqmm_sys_fill(self.h, rid, kind, seed, tag, _p(p, C.c_double), len(p), _p(v, C.c_double)),
              "cqmm_sys_fill")
        return v


class DeviceArray:
    """float64 buffer in HBM allocated through the library (no torch needed)."""

    def __init__(self, n):
        self.n = int(n)
        p = C.c_void_p()
        check(lib().cqmm_device_alloc(C.byref(p), max(self.n, 1) * 8), "cqmm_device_alloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        d = cls(arr.size)
        if arr.size:
            check(lib().cqmm_memcpy_h2d(d.ptr, arr.ctypes.data, arr.size * 8), "cqmm_memcpy_h2d")
        return d

    def to_host(self, out=None):
        if out is None:
            out = np.empty(self.n, np.float64)
        if self.n:
            check(lib().cqmm_memcpy_d2h(out.ctypes.data, self.ptr, self.n * 8), "cqmm_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().cqmm_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib().cqmm_synchronize(), "cqmm_synchronize")


def device_copy(dst, src, n=None):
    """dst[:n] = src[:n] on the library stream (streaming copy kernel)."""
    n = min(dst.n, src.n) if n is None else int(n)
    check(lib().cqmm_device_copy(dst.ptr, src.ptr, n), "cqmm_device_copy")


def timer_start():
    check(lib().cqmm_timer_start(), "cqmm_timer_start")


def timer_stop():
    """Milliseconds since timer_start(), measured with HIP events on the library stream."""
    ms = C.c_float()
    check(lib().cqmm_timer_stop(C.byref(ms)), "cqmm_timer_stop")
    return ms.value


def init(device=-1):
    check(lib().cqmm_init(device), "cqmm_init")


class Plan:
    """A triplet plan for one product (one CONQUEST mult(x) entry)."""

    def __init__(self, desc: MultDesc, pairs=None):
        """pairs = (p, q, n1, n2): the result is symmetric with these transpose
        pairs of its blocks (cqmm_plan_create_symmetric); None: every block."""
        init()
        self.desc = desc
        h = C.c_void_p()
        if pairs is None:
            check(lib().cqmm_plan_create(C.byref(desc), C.byref(h)), "cqmm_plan_create")
        else:
            p, q, n1, n2 = (np.ascontiguousarray(x, t) for x, t in
                            zip(pairs, (np.int64, np.int64, np.int32, np.int32)))
            self._pairs = (p, q, n1, n2)
            ptr = lambda x, t: x.ctypes.data_as(C.POINTER(t))
            check(lib().cqmm_plan_create_symmetric(C.byref(desc), len(p), ptr(p, C.c_int64), ptr(q, C.c_int64),
                                                   ptr(n1, C.c_int32), ptr(n2, C.c_int32), C.byref(h)),
                  "cqmm_plan_create_symmetric")
        self.h = h
        self.mult_type = desc.mult_type
        self.lena, self.lenb, self.lenc = desc.lena, desc.lenb, desc.lenc

    def close(self):
        if self.h:
            lib().cqmm_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = PlanInfo()
        check(lib().cqmm_plan_get_info(self.h, C.byref(i)), "cqmm_plan_get_info")
        return i

    def set_kernel(self, kernel):
        """KERNEL_AUTO, KERNEL_TRIPLET or KERNEL_TILE for the FP64 products."""
        check(lib().cqmm_plan_set_kernel(self.h, int(kernel)), "cqmm_plan_set_kernel")

    def set_tile_workgroups(self, wgs):
        """Class-tile kernel grid: 0 one workgroup per tile, else a persistent grid of wgs workgroups."""
        check(lib().cqmm_plan_set_tile_workgroups(self.h, int(wgs)), "cqmm_plan_set_tile_workgroups")

    def tune_dev(self, a_ptr, lena, b_ptr, lenb, c_ptr, lenc, flags=CQMM_DEVICE_PTRS, reps=2):
        """Time the plan's kernels on these device operands and keep the faster;
        returns (ms_tile or None, ms_triplet)."""
        mt, mp = C.c_float(), C.c_float()
        check(lib().cqmm_plan_tune(self.h, a_ptr, lena, b_ptr, lenb, c_ptr, lenc, flags | CQMM_DEVICE_PTRS, reps,
                                   C.byref(mt), C.byref(mp)), "cqmm_plan_tune")
        return (mt.value if mt.value >= 0 else None), mp.value

    def triplets(self):
        n = self.info().n_triplets
        o = np.zeros(n, np.int64)
        p = np.zeros(n, np.int64)
        q = np.zeros(n, np.int64)
        check(lib().cqmm_plan_get_triplets(self.h, _p(o, C.c_int64), _p(p, C.c_int64), _p(q, C.c_int64)),
              "cqmm_plan_get_triplets")
        return o, p, q

    def mat_mult(self, a, b, c, flags=0):
        """Host arrays (numpy float64). Type 1 overwrites c; type 2 overwrites a."""
        for name, arr in (("a", a), ("b", b), ("c", c)):
            if arr.dtype != np.float64 or not arr.flags.c_contiguous:
                raise CqmmError(f"mat_mult: {name} must be a contiguous float64 array")
        check(lib().cqmm_mat_mult(self.h, a.ctypes.data, a.size, b.ctypes.data, b.size, c.ctypes.data, c.size,
                                  flags), "cqmm_mat_mult")

    def mat_mult_dev(self, a_ptr, lena, b_ptr, lenb, c_ptr, lenc, flags=CQMM_DEVICE_PTRS):
        """Device pointers (e.g. torch tensors' data_ptr())."""
        check(lib().cqmm_mat_mult(self.h, a_ptr, lena, b_ptr, lenb, c_ptr, lenc, flags | CQMM_DEVICE_PTRS),
              "cqmm_mat_mult")

    def mat_mult_fp32_noalloc_dev(self, a_ptr, lena, b_ptr, lenb, c_ptr, lenc, flags=0):
        """cqmm_mat_mult's FP32 products on device pointers with the FP32 copies
        in the library's static scratch array (no device allocation)."""
        check(lib().cqmm_mat_mult_fp32_noalloc(self.h, a_ptr, lena, b_ptr, lenb, c_ptr, lenc,
                                               flags | CQMM_DEVICE_PTRS), "cqmm_mat_mult_fp32_noalloc")
