"""ctypes binding of libprofess_amd.so (the C ABI declared in include/ofd.h).

The library is required: importing this module raises when it is missing, and
every wrapper raises ``OFDError`` with the library's message on a nonzero status.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libprofess_amd.so")


class OFDError(RuntimeError):
    pass


class ofd_grid(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int), ("cell", C.c_double * 9)]


class ofd_functional(C.Structure):
    _fields_ = [("kinetic", C.c_int), ("lambda_", C.c_double), ("mu", C.c_double), ("alpha", C.c_double),
                ("beta", C.c_double), ("gamma", C.c_double), ("rho0", C.c_double), ("rhoS", C.c_double),
                ("precision", C.c_int)]


class ofd_opt_params(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("pot_tol", C.c_double), ("tole", C.c_double), ("conv_check", C.c_int),
                ("max_line", C.c_int), ("mixed_switch", C.c_double), ("fused_dots", C.c_int),
                ("cg_alg", C.c_int)]


class ofd_tune_result(C.Structure):
    _fields_ = [("rows_x", C.c_int), ("kb_y", C.c_int), ("kb_z", C.c_int), ("plane", C.c_int),
                ("configs_tried", C.c_int), ("us_default", C.c_double), ("us_best", C.c_double)]


class ofd_name(C.Structure):
    _fields_ = [("s", C.c_char * 16)]


class ofd_path(C.Structure):
    _fields_ = [("s", C.c_char * 256)]


class ofd_geometry(C.Structure):
    _fields_ = [("cell", C.c_double * 9), ("nion", C.c_int), ("nspecies", C.c_int),
                ("type_start", C.POINTER(C.c_int)), ("frac", C.POINTER(C.c_double)),
                ("file_index", C.POINTER(C.c_int)), ("species", C.POINTER(ofd_name)),
                ("pseudo_file", C.POINTER(ofd_path))]


class ofd_ion_params(C.Structure):
    _fields_ = [("cg_type", C.c_int), ("force_cutoff", C.c_double), ("max_steps", C.c_int), ("nn_dist", C.c_double),
                ("rho_dir_type", C.c_int)]


class ofd_ion_stats(C.Structure):
    _fields_ = [("steps", C.c_int), ("geometries", C.c_int), ("max_force", C.c_double), ("seconds", C.c_double)]


class ofd_spin(C.Structure):
    _fields_ = [("nspin", C.c_int), ("lsda", C.c_int), ("fixed_mag", C.c_int), ("nmag", C.c_int),
                ("tolm", C.c_double)]


class ofd_opt_stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("num_evals", C.c_int), ("newton_iters_total", C.c_int),
                ("fp32_evals", C.c_int), ("final_pot_norm", C.c_double), ("seconds", C.c_double)]


_lib = None

# every symbol the header declares (checked by the CPU test suite)
EXPORTS = ["ofd_create", "ofd_destroy", "ofd_last_error", "ofd_get_params", "ofd_set_ion_terms",
           "ofd_build_ion_potential", "ofd_get_ion_potential", "ofd_set_wgc_table", "ofd_build_wgc_table",
           "ofd_wgc_kernel_table",
           "ofd_calpot", "ofd_calpot_device", "ofd_rho_opt", "ofd_ion_forces", "ofd_pool_create",
           "ofd_pool_destroy", "ofd_pool_last_error", "ofd_pool_admit", "ofd_pool_step", "ofd_pool_collect",
           "ofd_request_bytes", "ofd_calpot_enqueue", "ofd_sums_to_etable",
           "ofd_calpot_noalloc", "ofd_noalloc_limit_bytes", "ofd_get_stream", "ofd_sync",
           "ofd_calpot_many", "ofd_autotune", "ofd_pools_step", "ofd_pool_finished", "ofd_server_create",
           "ofd_server_add_pool", "ofd_server_wait", "ofd_server_last_error", "ofd_server_destroy",
           "ofd_get_volume", "ofd_ewald", "ofd_ion_relax", "ofd_read_geometry", "ofd_free_geometry",
           "ofd_read_recpot", "ofd_write_density", "ofd_read_density", "ofd_set_spin", "ofd_get_magnetization",
           "ofd_spin_split"]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OFDError("libprofess_amd.so is not built (run `python -m profess_amd.build`)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    L.ofd_create.argtypes = [C.POINTER(ofd_grid), C.POINTER(ofd_functional), C.c_double, C.POINTER(vp)]
    L.ofd_destroy.argtypes = [vp]
    L.ofd_destroy.restype = None
    L.ofd_last_error.argtypes = [vp]
    L.ofd_last_error.restype = C.c_char_p
    L.ofd_get_params.argtypes = [vp, dp]
    L.ofd_get_volume.argtypes = [vp, dp]
    L.ofd_set_ion_terms.argtypes = [vp, dp, C.c_double]
    L.ofd_build_ion_potential.argtypes = [vp, C.c_int, ip, dp, dp, ip, ip, dp, C.c_int, C.c_double]
    L.ofd_get_ion_potential.argtypes = [vp, dp]
    L.ofd_set_wgc_table.argtypes = [vp, C.c_int, dp, dp, dp, dp]
    L.ofd_build_wgc_table.argtypes = [vp, ip]
    L.ofd_wgc_kernel_table.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, dp, dp]
    L.ofd_calpot.argtypes = [vp, dp, C.c_int, dp, dp]
    L.ofd_calpot_device.argtypes = [vp, C.c_void_p, C.c_int, C.c_void_p, dp]
    L.ofd_rho_opt.argtypes = [vp, dp, C.c_int, C.POINTER(ofd_opt_params), dp, ip, C.POINTER(ofd_opt_stats)]
    L.ofd_ion_forces.argtypes = [vp, dp, dp]
    L.ofd_ewald.argtypes = [vp, dp, dp, dp]
    L.ofd_read_geometry.argtypes = [C.c_char_p, C.POINTER(ofd_geometry), C.c_char_p, C.c_int]
    L.ofd_free_geometry.argtypes = [C.POINTER(ofd_geometry)]
    L.ofd_free_geometry.restype = None
    L.ofd_read_recpot.argtypes = [C.c_char_p, dp, dp, C.c_int, ip, C.c_char_p, C.c_int]
    L.ofd_write_density.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_char_p, C.c_int]
    L.ofd_read_density.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, dp, C.c_double, C.c_int, C.c_int,
                                   C.c_int, dp, C.c_char_p, C.c_int]
    L.ofd_ion_relax.argtypes = [vp, dp, dp, C.POINTER(ofd_opt_params), C.POINTER(ofd_ion_params), dp, dp, dp,
                                C.c_int, C.POINTER(ofd_ion_stats)]
    L.ofd_pool_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
    L.ofd_pool_destroy.argtypes = [vp]
    L.ofd_pool_destroy.restype = None
    L.ofd_pool_last_error.argtypes = [vp]
    L.ofd_pool_last_error.restype = C.c_char_p
    L.ofd_pool_admit.argtypes = [vp, vp, dp, C.POINTER(ofd_opt_params)]
    L.ofd_pool_step.argtypes = [vp]
    L.ofd_pool_collect.argtypes = [vp, C.c_int, dp, dp, ip, C.POINTER(ofd_opt_stats), C.POINTER(vp)]
    L.ofd_request_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.ofd_request_bytes.restype = C.c_longlong
    L.ofd_calpot_enqueue.argtypes = [vp, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    L.ofd_sums_to_etable.argtypes = [vp, dp, dp]
    L.ofd_calpot_noalloc.argtypes = [vp, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    L.ofd_noalloc_limit_bytes.argtypes = []
    L.ofd_noalloc_limit_bytes.restype = C.c_longlong
    L.ofd_get_stream.argtypes = [vp]
    L.ofd_get_stream.restype = C.c_void_p
    L.ofd_sync.argtypes = [vp]
    L.ofd_calpot_many.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int, dp]
    L.ofd_autotune.argtypes = [vp, C.c_int, C.POINTER(ofd_tune_result)]
    L.ofd_pools_step.argtypes = [C.POINTER(vp), C.c_int]
    L.ofd_pool_finished.argtypes = [vp, ip, C.c_int]
    L.ofd_server_create.argtypes = [C.POINTER(vp)]
    L.ofd_server_add_pool.argtypes = [vp, vp]
    L.ofd_server_wait.argtypes = [vp, C.c_double]
    L.ofd_server_last_error.argtypes = [vp]
    L.ofd_server_last_error.restype = C.c_char_p
    L.ofd_server_destroy.argtypes = [vp]
    L.ofd_server_destroy.restype = None
    L.ofd_set_spin.argtypes = [vp, C.POINTER(ofd_spin)]
    L.ofd_get_magnetization.argtypes = [vp, dp, ip]
    L.ofd_spin_split.argtypes = [vp, dp, C.c_double, dp]
    _lib = L
    return L


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_int))
