"""ctypes binding of the C ABI in include/dsmc_abi.h (libcudadsmc.so).

The library is the product: every simulation call goes through it.  If it
is missing the import fails loudly -- there is no Python or CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcudadsmc.so")

DSMC_OK = 0
DSMC_ERR_ARG, DSMC_ERR_HIP, DSMC_ERR_CAPACITY, DSMC_ERR_GEOMETRY, DSMC_ERR_NODEVICE = -1, -2, -3, -4, -5
BC_STREAM, BC_SPECULAR, BC_DIFFUSE = 0, 1, 2
DEV_ERR_BOUNCE, DEV_ERR_INFLOW, DEV_ERR_CAPACITY, DEV_ERR_FILL = 1, 2, 4, 8


class dsmc_params(C.Structure):
    _fields_ = [
        ("boltz", C.c_double), ("mass", C.c_double), ("diam", C.c_double),
        ("Vtl", C.c_double), ("Tref", C.c_double),
        ("L1", C.c_double), ("L2", C.c_double), ("L3", C.c_double),
        ("N1", C.c_int), ("N2", C.c_int),
        ("n_inf", C.c_double), ("T", C.c_double), ("V_jet", C.c_double), ("V_y", C.c_double),
        ("T_wall", C.c_double), ("wall_bc", C.c_int),
        ("face_bc", C.c_int * 4), ("face_T", C.c_double * 4),
        ("tau", C.c_double), ("ppc", C.c_double), ("Fn", C.c_double),
        ("seed", C.c_uint64), ("capacity", C.c_int64), ("sample_start", C.c_int),
        ("zeta", C.c_int), ("rot_prob", C.c_double),
    ]


class dsmc_body(C.Structure):
    _fields_ = [("n_points", C.c_int), ("xy", C.POINTER(C.c_double))]


class dsmc_case_info(C.Structure):
    _fields_ = [
        ("ncell", C.c_int), ("nseg", C.c_int), ("ncut", C.c_int), ("cseg_len", C.c_int),
        ("Fn", C.c_double), ("dt", C.c_double),
        ("L1", C.c_float), ("L2", C.c_float), ("dx", C.c_float), ("dy", C.c_float),
        ("ntc_coef", C.c_float), ("K1", C.c_float), ("e_exp", C.c_float),
        ("sigcmax0", C.c_float),
        ("body_vmp", C.c_float), ("body_sig", C.c_float),
        ("face_vmp", C.c_float * 4), ("face_sig", C.c_float * 4),
        ("inj_e", C.c_double * 4), ("inj_a", C.c_float * 4),
        ("inj_vmp", C.c_float), ("inj_sig", C.c_float), ("inj_ux", C.c_float), ("inj_uy", C.c_float),
        ("margin", C.c_float),
        ("body_bc", C.c_int), ("face_bc", C.c_int * 4),
        ("seed", C.c_uint32),
        ("cell_volume", C.c_double),
        ("ppc_eff", C.c_double),
        ("capacity", C.c_int64),
        ("zeta", C.c_int), ("sample_blocks", C.c_int),
        ("rot_p", C.c_float), ("lb_ia", C.c_float), ("lb_ib", C.c_float),
        ("body_ekt", C.c_float), ("face_ekt", C.c_float * 4), ("inj_ekt", C.c_float),
    ]


class dsmc_status(C.Structure):
    _fields_ = [
        ("step", C.c_uint32), ("n_live", C.c_uint32), ("next_pid", C.c_uint32), ("err", C.c_uint32),
        ("particle_steps", C.c_uint64), ("collisions", C.c_uint64), ("nsamples", C.c_uint64),
        ("last_collisions", C.c_uint32), ("n_move", C.c_uint32),
    ]


class dsmc_launch(C.Structure):
    _fields_ = [("move_blocks", C.c_int), ("scatter_blocks", C.c_int), ("cell_blocks", C.c_int),
                ("cell_tile", C.c_int), ("graph_steps", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class dsmc_grid(C.Structure):
    _fields_ = [("N1", C.c_int), ("N2", C.c_int), ("L1", C.c_double), ("L2", C.c_double)]


class dsmc_cut_tables(C.Structure):
    _fields_ = [("nseg", C.c_int), ("seg", C.POINTER(C.c_double)), ("frac", C.POINTER(C.c_double)),
                ("cut", C.POINTER(C.c_uint8))]


_P = C.POINTER
_f32p = _P(C.c_float)
_f64p = _P(C.c_double)
_i32p = _P(C.c_int)
_u32p = _P(C.c_uint32)
_u8p = _P(C.c_uint8)
_u64p = _P(C.c_uint64)
_h = C.c_void_p

# name -> argtypes; every function returns int status except dsmc_last_error
SIGNATURES = {
    "dsmc_device_count": [_i32p],
    "dsmc_device_cus": [_i32p],
    "dsmc_create": [_P(dsmc_params), _P(dsmc_body), _P(C.c_void_p)],
    "dsmc_create_from_tables": [_P(dsmc_params), _P(dsmc_cut_tables), _P(C.c_void_p)],
    "dsmc_destroy": [_h],
    "dsmc_case_info_get": [_h, _P(dsmc_case_info)],
    "dsmc_case_tables": [_h, _f32p, _i32p, _i32p, _f64p, _f32p, _u8p, _f32p, _u64p],
    "dsmc_case_describe": [_P(dsmc_params), _P(dsmc_body), _P(dsmc_case_info), _f32p, _i32p, _i32p,
                           _f64p, _f32p, _u8p, _f64p, _f32p, _u64p],
    "dsmc_init_freestream": [_h],
    "dsmc_init_fields": [_h, _f64p, _f64p, _f64p, _f64p],
    "dsmc_upload_particles": [_h, C.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p, _u32p],
    "dsmc_download_particles": [_h, C.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p, _u32p, _P(C.c_int64)],
    "dsmc_move": [_h],
    "dsmc_sample": [_h],
    "dsmc_bin_sort": [_P(dsmc_grid), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                      C.c_void_p, C.c_void_p],
    "dsmc_bin_sort_noalloc": [_P(dsmc_grid), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                              C.c_void_p, C.c_void_p, C.c_void_p],
    "dsmc_reassign": [_h],
    "dsmc_collide": [_h],
    "dsmc_run": [_h, C.c_int, C.c_int],
    "dsmc_sync": [_h],
    "dsmc_status_get": [_h, _P(dsmc_status)],
    "dsmc_set_keep_perm": [_h, C.c_int],
    "dsmc_set_subcells": [_h, _u32p],
    "dsmc_amr_update": [_h],
    "dsmc_download_amr": [_h, _u32p, _f32p, C.POINTER(C.c_double)],
    "dsmc_download_moved": [_h, C.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p, _u32p, _P(C.c_int64)],
    "dsmc_download_bins": [_h, C.c_int64, _i32p, _i32p],
    "dsmc_download_start": [_h, _i32p],
    "dsmc_download_perm": [_h, C.c_int64, _i32p],
    "dsmc_cell_state": [_h, _f32p, _f32p, _i32p],
    "dsmc_set_cell_state": [_h, _f32p, _f32p],
    "dsmc_restore_counters": [_h, _P(dsmc_status)],
    "dsmc_upload_samples": [_h, _f64p, C.c_uint64],
    "dsmc_reset_samples": [_h],
    "dsmc_download_samples": [_h, _f64p, _P(C.c_uint64)],
    "dsmc_download_fields": [_h, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p],
    "dsmc_download_rot_temperature": [_h, _f64p],
    "dsmc_upload_rotational": [_h, C.c_int64, _f32p],
    "dsmc_download_rotational": [_h, C.c_int, C.c_int64, _f32p, _P(C.c_int64)],
    "dsmc_batch_create": [_P(dsmc_params), _P(dsmc_body), C.c_int, _P(C.c_void_p)],
    "dsmc_batch_destroy": [_h],
    "dsmc_batch_case_info": [_h, C.c_int, _P(dsmc_case_info)],
    "dsmc_batch_init_freestream": [_h],
    "dsmc_batch_init_fields": [_h, _f64p, _f64p, _f64p, _f64p],
    "dsmc_batch_run": [_h, C.c_int, C.c_int],
    "dsmc_batch_sync": [_h],
    "dsmc_batch_status": [_h, C.c_int, _P(dsmc_status)],
    "dsmc_batch_download_particles": [_h, C.c_int, C.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p, _u32p,
                                      _P(C.c_int64)],
    "dsmc_batch_cell_state": [_h, C.c_int, _f32p, _f32p, _i32p],
    "dsmc_batch_reset_samples": [_h],
    "dsmc_batch_download_samples": [_h, C.c_int, _f64p, _P(C.c_uint64)],
    "dsmc_batch_download_fields": [_h, C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p],
    "dsmc_batch_download_counts": [_h, _f64p, _P(C.c_uint64)],
    "dsmc_batch_download_moments": [_h, _f64p, _P(C.c_uint64)],
    "dsmc_batch_wall_distance": [_h, _f64p],
    "dsmc_batch_compact": [_h, _i32p, C.c_int],
    "dsmc_batch_set_timing": [_h, C.c_int],
    "dsmc_batch_kernel_times": [_h, _f64p, _f64p, _P(C.c_int64)],
    "dsmc_get_launch": [_h, _P(dsmc_launch)],
    "dsmc_set_launch": [_h, _P(dsmc_launch)],
    "dsmc_batch_get_launch": [_h, _P(dsmc_launch)],
    "dsmc_batch_set_launch": [_h, _P(dsmc_launch)],
    "dsmc_set_timing": [_h, C.c_int],
    "dsmc_kernel_times": [_h, _f64p, _f64p, _P(C.c_int64)],
}

_lib = None


def load(path=LIB_PATH):
    """Loads libcudadsmc.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.dsmc_last_error.argtypes = []
    lib.dsmc_last_error.restype = C.c_char_p
    _lib = lib
    return lib


class DSMCError(RuntimeError):
    pass


def check(rc):
    if rc != DSMC_OK:
        msg = load().dsmc_last_error().decode(errors="replace")
        raise DSMCError(f"dsmc status {rc}: {msg}")


def ptr(a, ctype):
    """ctypes pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(_P(ctype))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)
