"""ctypes binding of libbbgpu.so (the C-ABI declared in include/bbgpu.h).

The product path has no CPU fallback: if the library is missing, or no GPU is visible when a kernel is
called, these functions raise.
"""

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libbbgpu.so")
MAX_LINES = 8


class BBError(RuntimeError):
    """An error reported by libbbgpu (bb_last_error)."""


class bb_grid(C.Structure):
    _fields_ = [("nz", C.c_int32), ("ltype", C.c_int32), ("axtyp", C.c_int32), ("velsys", C.c_int32),
                ("nlines", C.c_int32), ("nzq", C.c_int32), ("cmode", C.c_int32), ("pad0", C.c_int32),
                ("radsep", C.c_double), ("cdens", C.c_double), ("pixarea", C.c_double), ("cdelt0", C.c_double),
                ("cdelt1", C.c_double), ("crota2", C.c_double), ("sig_instr", C.c_double), ("crval3", C.c_double),
                ("cdelt3", C.c_double), ("crpix3", C.c_double), ("chmin", C.c_double), ("zstep", C.c_double),
                ("relint", C.c_double * MAX_LINES), ("lfreq0", C.c_double * MAX_LINES)]


class bb_box(C.Structure):
    _fields_ = [("blo0", C.c_int32), ("blo1", C.c_int32), ("bhi0", C.c_int32), ("bhi1", C.c_int32)]


class bb_profile_desc(C.Structure):
    _fields_ = [("nr", C.c_int32), ("npix", C.c_int32), ("nok", C.c_int32), ("pad0", C.c_int32),
                ("smooth", C.c_double), ("corr", C.c_double), ("floor", C.c_double), ("norm", C.c_double),
                ("dmed", C.c_double)]


NKEY = 21      # keys of the likelihood context (bbgpu.h BB_K_*)


class bb_ring_plan(C.Structure):
    _fields_ = [("nmr", C.c_int32), ("nsub", C.c_int32), ("thick", C.c_int32), ("from_fit", C.c_uint32),
                ("ridx", C.c_void_p), ("rad", C.c_void_p), ("base", C.c_void_p), ("idx", C.c_void_p),
                ("dr", C.c_void_p), ("cst", C.c_void_p), ("rads", C.c_void_p), ("densw", C.c_void_p),
                ("side", C.c_void_p)]


class bb_like_desc(C.Structure):
    _fields_ = [("nr", C.c_int32), ("ndim", C.c_int32), ("chunk", C.c_int32), ("mc", C.c_int32), ("nvpar", C.c_int32),
                ("seed", C.c_uint32), ("box_free", C.c_int32), ("halves", C.c_int32), ("likebox", bb_box),
                ("fixbox", bb_box), ("sigma", C.c_double), ("likefactor", C.c_double), ("eps", C.c_double),
                ("ext", C.c_double), ("pixas0", C.c_double), ("pixas1", C.c_double), ("nx", C.c_int32),
                ("ny", C.c_int32), ("mc_cap", C.c_int64), ("edgewidth", C.c_double), ("oedge", C.c_double),
                ("edge0", C.c_double), ("edgea0", C.c_double), ("edger0", C.c_double), ("annw", C.c_double),
                ("ann_cap", C.c_int32), ("empty", C.c_int32)]


class bb_like_profile(C.Structure):
    _fields_ = [("d", bb_profile_desc), ("wpix", C.c_void_p), ("ypix", C.c_void_p), ("lcoef", C.c_void_p),
                ("lcol", C.c_void_p), ("near", C.c_void_p), ("widx", C.c_void_p), ("wval", C.c_void_p),
                ("widx_e", C.c_void_p), ("wval_e", C.c_void_p), ("qunit", C.c_double), ("unit", C.c_double),
                ("radii", C.c_void_p), ("dmin", C.c_double)]


class bb_norm_par(C.Structure):
    _fields_ = [("side", C.c_int32), ("ftype", C.c_int32), ("normtype", C.c_int32), ("foot", C.c_int32),
                ("wpow", C.c_double), ("data_noise", C.c_double), ("bweight", C.c_double), ("hmax", C.c_double)]


_P = C.c_void_p
_SIGS = {
    "bb_last_error": (C.c_char_p, []),
    "bb_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "bb_model_create": (C.c_int, [C.POINTER(bb_grid), _P, _P, _P, C.c_int, C.c_int, C.c_double, C.POINTER(_P)]),
    "bb_model_destroy": (C.c_int, [_P]),
    "bb_model_set_tile": (C.c_int, [_P, C.c_int, C.c_int]),
    "bb_model_get_tile": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bb_model_set_bin": (C.c_int, [_P, C.c_int]),
    "bb_model_set_deposit": (C.c_int, [_P, C.c_int]),
    "bb_azplan_create": (C.c_int, [_P, _P, C.c_int, C.POINTER(_P)]),
    "bb_azplan_destroy": (C.c_int, [_P]),
    "bb_analytic_work_size": (C.c_int64, [_P, C.c_int]),
    "bb_galmod_analytic": (C.c_int, [_P, _P, bb_box, _P, C.c_int, C.c_int, _P, C.c_int64, _P, C.c_int, _P]),
    "bb_galmod_analytic_boxes": (C.c_int, [_P, _P, bb_box, _P, _P, C.c_int, C.c_int, _P, C.c_int64, _P, C.c_int,
                                           _P]),
    "bb_galmod_mc_boxes": (C.c_int, [_P, bb_box, _P, _P, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int64, _P,
                                     C.c_int64, _P, _P, _P, C.c_int, _P]),
    "bb_trace_deposit": (C.c_int, [_P, _P]),
    "bb_conv2d_same": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, _P]),
    "bb_residual_work_size": (C.c_int64, [C.c_int, C.c_int64]),
    "bb_residual_gauss": (C.c_int, [_P, _P, C.c_int, C.c_int64, C.c_double, C.c_double, _P, C.c_int64, _P, _P,
                                    _P]),
    "bb_conv_residual_work_size": (C.c_int64, [_P, C.c_int, C.c_int, C.c_int]),
    "bb_dilate_mask": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _P, _P, _P]),
    "bb_dilate_rmax": (C.c_int, []),
    "bb_noise_flags": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P]),
    "bb_masked_median_work_size": (C.c_int64, []),
    "bb_masked_median": (C.c_int, [_P, _P, C.c_int64, _P, _P, _P, _P, _P]),
    "bb_masked_std_work_size": (C.c_int64, []),
    "bb_masked_std": (C.c_int, [_P, _P, C.c_int64, _P, _P, _P, _P]),
    "bb_flag_subsample_work_size": (C.c_int64, [C.c_int64, C.c_int64]),
    "bb_flag_subsample": (C.c_int, [_P, C.c_int64, C.c_int64, _P, C.c_int64, _P, C.c_int64, _P, _P]),
    "bb_box_dilate": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P]),
    "bb_like_maps": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, bb_box, _P, _P, _P, _P, _P]),
    "bb_lag_xy_max": (C.c_int, []),
    "bb_lag_xy_work_size": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "bb_lag_xy": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _P, C.c_int64, _P, _P, _P]),
    "bb_lag_z_work_size": (C.c_int64, [C.c_int, C.c_int, C.c_int]),
    "bb_lag_z": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _P, C.c_int64, _P, _P,
                           _P]),
    "bb_conv_residual_gauss": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _P, C.c_int64,
                                         _P, _P, _P]),
    "bb_conv_residual_noalloc_max": (C.c_int64, []),
    "bb_conv_residual_gauss_noalloc": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _P,
                                                 _P, _P]),
    "bb_mc_work_size": (C.c_int64, [_P, bb_box, C.c_int, C.c_int, C.c_int64]),
    "bb_galmod_mc": (C.c_int, [_P, bb_box, _P, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int64, _P, C.c_int64, _P,
                               _P, _P, C.c_int, _P]),
    "bb_norm_work_size": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bb_norm3dfit": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P,
                               _P, C.POINTER(bb_norm_par), _P, _P, _P, C.c_int, _P, _P, C.c_int64, _P, _P, _P, _P]),
    "bb_local_norm": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P]),
    "bb_ring_maps_work_size": (C.c_int64, [_P, C.c_int]),
    "bb_ring_maps_deposit": (C.c_int, [_P, _P, bb_box, _P, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int, C.c_int,
                                       C.c_double, _P, C.c_int64, _P, _P]),
    "bb_ring_maps_finish": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _P, _P]),
    "bb_profile_create": (C.c_int, [C.POINTER(bb_profile_desc), _P, _P, _P, _P, _P, C.POINTER(_P)]),
    "bb_profile_destroy": (C.c_int, [_P]),
    "bb_profile_work_size": (C.c_int64, [_P, C.c_int]),
    "bb_profile_solve": (C.c_int, [_P, _P, C.c_int, _P, C.c_int64, _P, _P]),
    "bb_profile_set_dens": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, _P, _P, _P]),
    "bb_ctx_create": (C.c_int, [_P, C.POINTER(bb_like_desc), _P, _P, _P, _P, _P, _P, C.POINTER(bb_ring_plan),
                                C.POINTER(bb_ring_plan), C.POINTER(bb_like_profile), C.POINTER(_P)]),
    "bb_ctx_destroy": (C.c_int, [_P]),
    "bb_loglike_launch": (C.c_int, [_P, _P, C.c_int, _P]),
    "bb_loglike_collect": (C.c_int, [_P, _P, _P]),
    "bb_loglike_batch": (C.c_int, [_P, _P, C.c_int, _P, _P, _P]),
    "bb_ctx_tables": (C.c_int, [_P, _P, C.c_int, _P, _P]),
    "bb_ctx_info": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "bb_ctx_set_halves": (C.c_int, [_P, C.c_int]),
    "bb_ctx_set_graphs": (C.c_int, [_P, C.c_int]),
    "bb_ctx_graph_replays": (C.c_int64, [_P]),
    "bb_ctx_trace": (C.c_int, [_P, _P]),
    "bb_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_P)]),
    "bb_mlp_destroy": (C.c_int, [_P]),
    "bb_mlp_param_count": (C.c_int64, [_P]),
    "bb_mlp_set_params": (C.c_int, [_P, _P]),
    "bb_mlp_get_params": (C.c_int, [_P, _P]),
    "bb_mlp_forward": (C.c_int, [_P, _P, C.c_int, _P, _P]),
    "bb_mlp_grad": (C.c_int, [_P, _P, _P, C.c_int, _P, _P, _P]),
    "bb_mlp_train": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_float, _P, _P]),
    "bb_mlp_walk": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P]),
}

_lib = None


def lib():
    """The loaded library (loads it on first use); raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIBPATH):
            raise BBError(f"{LIBPATH} is not built: run python -m bbarolo2_amd.build")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise BBError(lib().bb_last_error().decode() or f"libbbgpu error {rc}")


def declared_symbols():
    """The entry points declared in include/bbgpu.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "bbgpu.h")
    with open(hdr) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(bb_\w+)\s*\(", txt, re.M)))
