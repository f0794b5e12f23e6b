"""ctypes binding of liblaminocam.so (the gfx950 HIP library).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liblaminocam.so")

LC_DEVICE_PTRS = 0x1
LC_GRAPH = 0x2
LC_X0_ZERO = 0x4
LC_TV_ZERO_PAD = 0x8
LC_TV_ANISO = 0x10
LC_GEOM_LAZY_PSF = 0x20
LC_FP64 = 0x40
LC_REQ_BORROW = 0x80
LC_REQ_TEST_FAIL = 0x100
LC_HOST_SYNC = 0x200

LC_REG_SPLIT_BREGMAN = 0
LC_REG_UNCONSTRAINED = 1

# every symbol declared in include/laminocam.h
EXPORTS = [
    "lc_abi_version", "lc_last_error", "lc_ctx_create", "lc_ctx_destroy", "lc_ctx_sync",
    "lc_geom2d_create", "lc_geom2d_destroy", "lc_geom2d_info", "lc_psf2d_kernel_hat",
    "lc_sysmat2d_psf", "lc_forward2d", "lc_backproj2d", "lc_sysmat2d_nufft", "lc_cg2d", "lc_split_bregman2d", "lc_grad", "lc_divergence",
    "lc_laplacian", "lc_batch2d_create", "lc_batch2d_destroy", "lc_batch2d_set",
    "lc_batch2d_get", "lc_batch2d_device_ptr", "lc_batch2d_cg", "lc_batch2d_split_bregman",
    "lc_batch2d_reports", "lc_batch2d_sysmat", "lc_batch2d_profile",
    "lc_geom3d_create", "lc_geom3d_destroy", "lc_geom3d_info", "lc_psf3d_kernel_hat",
    "lc_server_create", "lc_server_submit", "lc_server_poll", "lc_server_wait",
    "lc_server_result", "lc_server_stats", "lc_server_timeline", "lc_server_destroy",
    "lc_sysmat3d_psf", "lc_forward3d", "lc_backproj3d", "lc_sysmat3d_nufft", "lc_cg3d", "lc_split_bregman3d", "lc_batch_create", "lc_autotune2d", "lc_set_tuning2d",
    "lc_get_tuning2d", "lc_autotune3d", "lc_set_tuning3d", "lc_get_tuning3d",
    "lc_autotune2d_nufft", "lc_set_tuning2d_nufft", "lc_get_tuning2d_nufft",
    "lc_sysmat2d_psf_f64", "lc_cg2d_f64", "lc_split_bregman2d_f64", "lc_sysmat3d_psf_f64",
    "lc_cg3d_f64", "lc_split_bregman3d_f64", "lc_grad_f64", "lc_divergence_f64",
    "lc_laplacian_f64", "lc_batch_create_ex", "lc_batch2d_set_f64", "lc_batch2d_get_f64",
    "lc_batch2d_device_ptr_f64", "lc_psf2d_kernel_hat_f64", "lc_psf3d_kernel_hat_f64",
    "lc_recon_params_default", "lc_prepare_sinogram", "lc_prepare_sinogram_f64", "lc_mbir2d",
    "lc_mbir2d_f64", "lc_mbir3d", "lc_mbir3d_f64", "lc_cg2d_noalloc",
    "lc_reference_memory_model", "lc_volume_fits", "lc_volume_fits_one_gpu", "lc_request_bytes",
    "lc_tiff_info", "lc_tiff_read", "lc_tiff_write", "lc_h5_shape", "lc_h5_read",
]


class LcReport(C.Structure):
    _fields_ = [("iters", C.c_int64), ("outer", C.c_int64), ("res", C.c_double),
                ("dx", C.c_double), ("stop", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self):
        return {"iters": self.iters, "outer": self.outer, "res": self.res,
                "dx": self.dx, "stop": self.stop}


class LcRequest(C.Structure):
    """lc_request of laminocam.h"""
    _fields_ = [("g2", C.c_void_p), ("g3", C.c_void_p), ("nslices", C.c_int32),
                ("solver", C.c_int32), ("max_iter", C.c_int32), ("outer", C.c_int32),
                ("inner", C.c_int32), ("tol", C.c_double), ("xtol", C.c_double),
                ("lam", C.c_double), ("mu", C.c_double), ("flags", C.c_uint32),
                ("sino", C.c_void_p), ("rhs", C.c_void_p), ("x0", C.c_void_p),
                ("x_out", C.c_void_p)]


class LcReconParams(C.Structure):
    """lc_recon_params of laminocam.h (ReconParams, include/recon_params.h:45-54)"""
    _fields_ = [("regularizer", C.c_int32), ("pad", C.c_int32), ("recon_dims", C.c_int64 * 3),
                ("max_iters", C.c_int64), ("inner_iters", C.c_int64), ("lam", C.c_float),
                ("mu", C.c_float), ("tol", C.c_float), ("xtol", C.c_float),
                ("pad_factor", C.c_float), ("pad2", C.c_float)]


class LcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"laminocam error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load liblaminocam.so (build it with laminocam_amd.build.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    f32p = C.POINTER(C.c_float)
    f64p = C.POINTER(C.c_double)
    vp = C.c_void_p
    i64, i32, u32 = C.c_int64, C.c_int32, C.c_uint32
    sig = {
        "lc_abi_version": (C.c_int, []),
        "lc_last_error": (C.c_char_p, []),
        "lc_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "lc_ctx_destroy": (C.c_int, [vp]),
        "lc_ctx_sync": (C.c_int, [vp]),
        "lc_geom2d_create": (C.c_int, [vp, f64p, i64, i64, i64, u32, C.POINTER(vp)]),
        "lc_geom2d_destroy": (C.c_int, [vp]),
        "lc_geom2d_info": (C.c_int, [vp, C.POINTER(i64)]),
        "lc_psf2d_kernel_hat": (C.c_int, [vp, vp, vp]),
        "lc_sysmat2d_psf": (C.c_int, [vp, vp, vp, i64, vp, u32]),
        "lc_forward2d": (C.c_int, [vp, vp, vp, i64, vp, u32]),
        "lc_backproj2d": (C.c_int, [vp, vp, vp, i64, vp, u32]),
        "lc_sysmat2d_nufft": (C.c_int, [vp, vp, vp, i64, vp, u32]),
        "lc_cg2d": (C.c_int, [vp, vp, vp, vp, i64, i64, C.c_double, C.c_double, u32,
                              C.POINTER(LcReport)]),
        "lc_cg2d_noalloc": (C.c_int, [vp, vp, vp, vp, i64, i64, C.c_double, C.c_double, u32,
                                      C.POINTER(LcReport)]),
        "lc_split_bregman2d": (C.c_int, [vp, vp, vp, vp, i64, C.c_double, C.c_double, i64,
                                         i64, C.c_double, C.c_double, u32,
                                         C.POINTER(LcReport)]),
        "lc_grad": (C.c_int, [vp, vp, i64, i64, i64, vp, u32]),
        "lc_divergence": (C.c_int, [vp, vp, i64, i64, i64, vp, u32]),
        "lc_laplacian": (C.c_int, [vp, vp, i64, i64, i64, vp, u32]),
        "lc_batch2d_create": (C.c_int, [vp, i32, C.POINTER(vp), C.POINTER(i32), C.POINTER(vp)]),
        "lc_batch2d_destroy": (C.c_int, [vp]),
        "lc_batch2d_set": (C.c_int, [vp, i32, i32, vp, u32]),
        "lc_batch2d_get": (C.c_int, [vp, i32, i32, vp, u32]),
        "lc_batch2d_device_ptr": (C.c_int, [vp, i32, i32, C.POINTER(vp)]),
        "lc_batch2d_cg": (C.c_int, [vp, C.POINTER(i32), C.c_double, C.c_double, u32]),
        "lc_batch2d_split_bregman": (C.c_int, [vp, f32p, f32p, i64, i64, C.c_double,
                                               C.c_double, u32]),
        "lc_batch2d_reports": (C.c_int, [vp, C.POINTER(LcReport)]),
        "lc_batch2d_sysmat": (C.c_int, [vp, i32, i32]),
        "lc_batch2d_profile": (C.c_int, [vp, i32, i32, f64p, C.POINTER(i32)]),
        "lc_geom3d_create": (C.c_int, [vp, f64p, f64p, f64p, i64, i64, i64, C.POINTER(i64), u32,
                                       C.POINTER(vp)]),
        "lc_geom3d_destroy": (C.c_int, [vp]),
        "lc_geom3d_info": (C.c_int, [vp, C.POINTER(i64)]),
        "lc_psf3d_kernel_hat": (C.c_int, [vp, vp, vp]),
        "lc_sysmat3d_psf": (C.c_int, [vp, vp, vp, vp, u32]),
        "lc_server_create": (C.c_int, [vp, C.c_uint64, i32, i32, C.POINTER(vp)]),
        "lc_server_submit": (C.c_int, [vp, C.POINTER(LcRequest), C.POINTER(C.c_uint64)]),
        "lc_server_poll": (C.c_int, [vp, C.c_uint64, C.POINTER(i32)]),
        "lc_server_wait": (C.c_int, [vp, C.c_uint64, C.c_double, C.POINTER(i32)]),
        "lc_server_result": (C.c_int, [vp, C.c_uint64, vp, C.POINTER(LcReport), f64p]),
        "lc_server_stats": (C.c_int, [vp, f64p]),
        "lc_server_timeline": (C.c_int, [vp, i32, f64p, C.POINTER(i32)]),
        "lc_server_destroy": (C.c_int, [vp]),
        "lc_forward3d": (C.c_int, [vp, vp, vp, vp, u32]),
        "lc_backproj3d": (C.c_int, [vp, vp, vp, vp, vp, u32]),
        "lc_sysmat3d_nufft": (C.c_int, [vp, vp, vp, vp, u32]),
        "lc_cg3d": (C.c_int, [vp, vp, vp, vp, i64, C.c_double, C.c_double, u32,
                              C.POINTER(LcReport)]),
        "lc_split_bregman3d": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_double, i64, i64,
                                         C.c_double, C.c_double, u32, C.POINTER(LcReport)]),
        "lc_batch_create": (C.c_int, [vp, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(i32),
                                      C.POINTER(vp)]),
        "lc_autotune2d": (C.c_int, [vp, vp, i32, C.POINTER(i32), C.POINTER(C.c_double)]),
        "lc_set_tuning2d": (C.c_int, [vp, i64, C.POINTER(i32)]),
        "lc_get_tuning2d": (C.c_int, [vp, i64, C.POINTER(i32)]),
        "lc_autotune3d": (C.c_int, [vp, vp, C.POINTER(i32), C.POINTER(C.c_double)]),
        "lc_autotune2d_nufft": (C.c_int, [vp, vp, i32, C.POINTER(i32), C.POINTER(C.c_double)]),
        "lc_set_tuning2d_nufft": (C.c_int, [vp, vp, i32]),
        "lc_get_tuning2d_nufft": (C.c_int, [vp, vp, C.POINTER(i32)]),
        "lc_set_tuning3d": (C.c_int, [vp, vp, i32]),
        "lc_get_tuning3d": (C.c_int, [vp, vp, C.POINTER(i32)]),
        "lc_sysmat2d_psf_f64": (C.c_int, [vp, vp, vp, i64, vp, u32]),
        "lc_cg2d_f64": (C.c_int, [vp, vp, vp, vp, i64, i64, C.c_double, C.c_double, u32,
                                  C.POINTER(LcReport)]),
        "lc_split_bregman2d_f64": (C.c_int, [vp, vp, vp, vp, i64, C.c_double, C.c_double, i64,
                                             i64, C.c_double, C.c_double, u32,
                                             C.POINTER(LcReport)]),
        "lc_sysmat3d_psf_f64": (C.c_int, [vp, vp, vp, vp, u32]),
        "lc_cg3d_f64": (C.c_int, [vp, vp, vp, vp, i64, C.c_double, C.c_double, u32,
                                  C.POINTER(LcReport)]),
        "lc_split_bregman3d_f64": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_double, i64, i64,
                                             C.c_double, C.c_double, u32, C.POINTER(LcReport)]),
        "lc_grad_f64": (C.c_int, [vp, vp, i64, i64, i64, vp, u32]),
        "lc_divergence_f64": (C.c_int, [vp, vp, i64, i64, i64, vp, u32]),
        "lc_laplacian_f64": (C.c_int, [vp, vp, i64, i64, i64, vp, u32]),
        "lc_batch_create_ex": (C.c_int, [vp, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(i32),
                                         u32, C.POINTER(vp)]),
        "lc_batch2d_set_f64": (C.c_int, [vp, i32, i32, vp, u32]),
        "lc_batch2d_get_f64": (C.c_int, [vp, i32, i32, vp, u32]),
        "lc_batch2d_device_ptr_f64": (C.c_int, [vp, i32, i32, C.POINTER(vp)]),
        "lc_psf2d_kernel_hat_f64": (C.c_int, [vp, vp, vp]),
        "lc_psf3d_kernel_hat_f64": (C.c_int, [vp, vp, vp]),
        "lc_recon_params_default": (None, [C.POINTER(LcReconParams)]),
        "lc_reference_memory_model": (C.c_int, [C.POINTER(i64), i32, i32, f64p]),
        "lc_volume_fits": (C.c_int, [C.POINTER(LcReconParams), i32, C.c_uint64, C.POINTER(i32)]),
        "lc_volume_fits_one_gpu": (C.c_int, [vp, C.POINTER(LcReconParams), i32, C.POINTER(i32)]),
        "lc_request_bytes": (C.c_int, [C.POINTER(LcRequest), C.POINTER(C.c_uint64)]),
        "lc_tiff_info": (C.c_int, [C.c_char_p, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]),
        "lc_tiff_read": (C.c_int, [C.c_char_p, vp, i64, i64, i64]),
        "lc_tiff_write": (C.c_int, [C.c_char_p, vp, i64, i64, i64]),
        "lc_h5_shape": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(i32), C.POINTER(i64)]),
        "lc_h5_read": (C.c_int, [C.c_char_p, C.c_char_p, vp, i64, i32]),
        "lc_prepare_sinogram": (C.c_int, [vp, i32, C.POINTER(vp), C.POINTER(i64), i64, i64, vp, vp,
                                          u32]),
        "lc_prepare_sinogram_f64": (C.c_int, [vp, i32, C.POINTER(vp), C.POINTER(i64), i64, i64, vp,
                                              vp, u32]),
        "lc_mbir2d": (C.c_int, [vp, vp, i64, i64, i64, f64p, C.POINTER(LcReconParams), vp,
                                C.POINTER(LcReport), u32]),
        "lc_mbir2d_f64": (C.c_int, [vp, vp, i64, i64, i64, f64p, C.POINTER(LcReconParams), vp,
                                    C.POINTER(LcReport), u32]),
        "lc_mbir3d": (C.c_int, [vp, i32, C.POINTER(vp), C.POINTER(i64), i64, i64, f64p, f64p,
                                f64p, f64p, C.POINTER(LcReconParams), vp, C.POINTER(i64),
                                C.POINTER(LcReport), u32]),
        "lc_mbir3d_f64": (C.c_int, [vp, i32, C.POINTER(vp), C.POINTER(i64), i64, i64, f64p, f64p,
                                    f64p, f64p, C.POINTER(LcReconParams), vp, C.POINTER(i64),
                                    C.POINTER(LcReport), u32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status):
    if status != 0:
        msg = _lib.lc_last_error().decode(errors="replace") if _lib else ""
        raise LcError(status, msg)
