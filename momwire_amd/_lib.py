"""ctypes binding of libmomwire_hip.so (the C ABI in include/momwire_hip.h).

The shared library is built in-tree by ``python -m momwire_amd.build``.  There
is no fallback: if the library is missing or fails to load, every solve raises
``AcceleratorUnavailable`` -- the port-solution path runs on the GPU or not at
all.  ctypes releases the GIL for the duration of each foreign call, so a long
factorisation does not block other Python threads (the property the
reference's accelerators give by releasing the GIL, reference
tests/test_gil_release.py:43-105).
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmomwire_hip.so")

MW_OK, MW_EINVAL, MW_ESINGULAR, MW_EHIP, MW_EABORT, MW_ERANGE = range(6)
TSMAX = 18  # MW_TSMAX
TB = 16     # MW_TB


class AcceleratorUnavailable(RuntimeError):
    """The HIP library could not be loaded (not built, or no usable GPU)."""


class SolveAborted(RuntimeError):
    """A solve stopped because its cancel token was tripped (reference
    src/momwire/_cancel.py:20-79 raises the same-named exception)."""


class MwMesh(ctypes.Structure):
    _fields_ = [
        ("n_seg", ctypes.c_int32),
        ("n_edge", ctypes.c_int32),
        ("n_basis", ctypes.c_int32),
        ("n_tile", ctypes.c_int32),
        ("n_se", ctypes.c_int32),
        ("degree", ctypes.c_int32),
        ("n_qp", ctypes.c_int32),
        ("ek", ctypes.c_int32),
        ("seg", ctypes.c_void_p),
        ("seg_edge", ctypes.c_void_p),
        ("edge", ctypes.c_void_p),
        ("edge_i", ctypes.c_void_p),
        ("se_edge", ctypes.c_void_p),
        ("se_tab", ctypes.c_void_p),
        ("band_poly", ctypes.c_void_p),
        ("bw_loc", ctypes.c_void_p),
        ("bw_slot", ctypes.c_void_p),
        ("tile_b0", ctypes.c_void_p),
        ("tile_seg", ctypes.c_void_p),
        ("tile_nseg", ctypes.c_void_p),
        ("gl", ctypes.c_void_p),
        ("ek_label", ctypes.c_void_p),
        ("ek_img_obs", ctypes.c_void_p),
        ("ek_img_src", ctypes.c_void_p),
        ("se_static", ctypes.c_void_p),
    ]


class MwPorts(ctypes.Structure):
    _fields_ = [
        ("n_ports", ctypes.c_int32), ("n_con", ctypes.c_int32),
        ("col_max", ctypes.c_int32), ("con_max", ctypes.c_int32),
        ("port_max", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("col_idx", ctypes.c_void_p), ("col_val", ctypes.c_void_p),
        ("con_idx", ctypes.c_void_p), ("con_val", ctypes.c_void_p),
        ("port_idx", ctypes.c_void_p), ("port_val", ctypes.c_void_p),
    ]


class MwSweepProb(ctypes.Structure):
    _fields_ = [
        ("mesh", ctypes.c_int32), ("symmetric", ctypes.c_int32),
        ("k", ctypes.c_double), ("omega", ctypes.c_double),
        ("A", ctypes.c_void_p), ("lda", ctypes.c_int64), ("n", ctypes.c_int64),
        ("ncols", ctypes.c_int64), ("se_tab", ctypes.c_void_p),
        ("tile_pairs", ctypes.c_void_p), ("n_pairs", ctypes.c_int32), ("n_se", ctypes.c_int32),
        ("ipiv", ctypes.c_void_p), ("work", ctypes.c_void_p), ("info", ctypes.c_void_p),
        ("Y", ctypes.c_void_p), ("S", ctypes.c_void_p), ("T", ctypes.c_void_p),
    ]


class MwMlp(ctypes.Structure):
    _fields_ = [("params", ctypes.c_void_p), ("grads", ctypes.c_void_p),
                ("adam_m", ctypes.c_void_p), ("adam_v", ctypes.c_void_p),
                ("w16", ctypes.c_void_p), ("wt16", ctypes.c_void_p)]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_D = ctypes.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "mw_last_error": (ctypes.c_char_p, []),
    "mw_abi_version": (ctypes.c_int, []),
    "mw_cancel_create": (ctypes.c_int, [ctypes.POINTER(ctypes.POINTER(ctypes.c_int32))]),
    "mw_cancel_destroy": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32)]),
    "mw_same_edge_table": (ctypes.c_int, [ctypes.POINTER(MwMesh), _D, _P]),
    "mw_same_edge_static": (ctypes.c_int, [ctypes.POINTER(MwMesh), _P, _P]),
    "mw_fill_z": (ctypes.c_int, [ctypes.POINTER(MwMesh), _D, _D, _D, _D, _P, _I64, _P, _I32,
                                 _I32, _P]),
    "mw_fill_z_far32": (ctypes.c_int, [ctypes.POINTER(MwMesh), _D, _D, _D, _D, _P, _I64, _P, _I32,
                                       _I32, _D, _P]),
    "mw_getrf_workspace_bytes": (_I64, [_I64]),
    "mw_sysv_workspace_bytes": (_I64, [_I64]),
    "mw_set_pivot_threshold": (ctypes.c_int, [_D]),
    "mw_get_pivot_threshold": (_D, []),
    "mw_zgetrf": (ctypes.c_int, [_I64, _P, _I64, _P, _P, _P, _P, _P]),
    "mw_cgetrf": (ctypes.c_int, [_I64, _P, _I64, _P, _P, _P, _P, _P]),
    "mw_zgesv": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _P, _P, _P, _P]),
    "mw_cgesv": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _P, _P, _P, _P]),
    "mw_zsysv": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _P, _P, _P, _P, _P]),
    "mw_csysv": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _P, _P, _P, _P, _P]),
    "mw_zgesv_noalloc": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _P, _P, _P]),
    "mw_zgetrs": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _P, _P, _I64, _P]),
    "mw_cgetrs": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _P, _P, _I64, _P]),
    "mw_zsytrs": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _P, _P, _I64, _P]),
    "mw_csytrs": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _P, _P, _I64, _P]),
    "mw_zgemm_sub": (ctypes.c_int, [_I64, _I64, _I64, _P, _I64, _P, _I64, _P, _I64, _P]),
    "mw_cgemm_sub": (ctypes.c_int, [_I64, _I64, _I64, _P, _I64, _P, _I64, _P, _I64, _P]),
    "mw_zgemm_tn_add": (ctypes.c_int, [_I64, _I64, _I64, _P, _I64, _P, _I64, _P, _I64, _P]),
    "mw_cgemm_tn_add": (ctypes.c_int, [_I64, _I64, _I64, _P, _I64, _P, _I64, _P, _I64, _P]),
    "mw_zsparse_rows": (ctypes.c_int, [_I64, _I32, _P, _P, _P, _I64, _I64, _P, _I64, _P]),
    "mw_zscatter_cols": (ctypes.c_int, [_I64, _I64, _I32, _P, _P, _P, _I64, _P]),
    "mw_zsmall_solve": (ctypes.c_int, [_I32, _I32, _P, _I64, _P, _I64, _P, _P]),
    "mw_sweep_plan_create": (ctypes.c_int, [_I32, _P, _P, _I32, _P, _I32, _D,
                                            ctypes.POINTER(ctypes.c_void_p)]),
    "mw_sweep_run": (ctypes.c_int, [_P, _D, _D, _P, _P]),
    "mw_sweep_fill": (ctypes.c_int, [_P, _D, _D, _P]),
    "mw_zentries": (ctypes.c_int, [ctypes.POINTER(MwMesh), _P, _P, _D, _D, _D, _D, _P, _P, _I64,
                                   _I32, _P, _P]),
    "mw_hm_fill_dense": (ctypes.c_int, [ctypes.POINTER(MwMesh), _P, _P, _D, _D, _D, _D, _P, _P,
                                        _P, _P, _I64, _I32, _P, _P]),
    "mw_hm_aca": (ctypes.c_int, [ctypes.POINTER(MwMesh), _P, _P, _D, _D, _D, _D, _P, _P, _P, _P,
                                 _I32, _I32, _D, _P, _P, _P]),
    "mw_hm_apply": (ctypes.c_int, [_P, _P, _P, _I32, _P, _P, _P, _I32, _P, _P, _P, _I64, _I32,
                                   _P, _P, _I64, _P]),
    "mw_hm_invert": (ctypes.c_int, [_P, _P, _I32, _P, _P, _P, _P]),
    "mw_sweep_plan_destroy": (ctypes.c_int, [_P]),
    "mw_somm_six_integrals": (ctypes.c_int, [_D, _D, _D, _P, _P, _I64, _D, _I32, _P, _P, _P]),
    "mw_somm_surfaces": (ctypes.c_int, [_D, _D, _D, _D, _D, _P, _P, _I64, _D, _P, _P, _P]),
    "mw_somm_eval": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P]),
    "mw_ground_fields": (ctypes.c_int, [_P, _D, _D, _I32, _P, _I32, _P, _P, _I64, _I32, _I32,
                                        _P, _I64, _P]),
    "mw_fill_ground": (ctypes.c_int, [ctypes.POINTER(MwMesh), _D, _D, _D, _D, _D, _D, _D, _P,
                                      _I32, _P, _I64, _P, _I32, _I32, _P]),
    "mw_fill_ground_refl": (ctypes.c_int, [ctypes.POINTER(MwMesh), _D, _D, _D, _D, _D, _D, _D, _D,
                                           _D, _D, _D, _P, _I64, _P, _I32, _I32, _P]),
    "mw_zadd_sparse": (ctypes.c_int, [_I64, _I64, _P, _P, _P, _P, _I64, _P]),
    "mw_element_moments": (ctypes.c_int, [_I32, _I32, _I32, _P, _P, _P, _P, _P, _P, _I64, _P,
                                          _I64, _P, _I64, _P]),
    "mw_far_moments": (ctypes.c_int, [_I32, _P, _P, _I64, _I32, _I32, _I32, _P, _P, _D, _I32, _D,
                                      _D, _D, _P, _P, _I64, _P]),
    "mw_near_fields": (ctypes.c_int, [_I32, _P, _I32, _P, _P, _I64, _I32, _P, _P, _I64, _D, _I32,
                                      _D, _D, _D, _I32, _I32, _P, _I64, _P]),
    "mw_somm_surfaces_varied": (ctypes.c_int, [_P, _D, _D, _D, _P, _P, _I64, _D, _P, _P, _P]),
    "mw_mlp_param_count": (_I64, []),
    "mw_mlp_w16_count": (_I64, []),
    "mw_mlp_workspace_bytes": (_I64, [_I64]),
    "mw_mlp_sync_weights": (ctypes.c_int, [ctypes.POINTER(MwMlp), _P]),
    "mw_mlp_forward": (ctypes.c_int, [ctypes.POINTER(MwMlp), _P, _I64, _P, _P, _P]),
    "mw_mlp_backward": (ctypes.c_int, [ctypes.POINTER(MwMlp), _P, _P, _I64, _P, _P, _P]),
    "mw_mlp_adam_step": (ctypes.c_int, [ctypes.POINTER(MwMlp), ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_float, _I32, _P]),
    "mw_mlp_grid_values": (ctypes.c_int, [ctypes.POINTER(MwMlp), _D, _D, _P, _P, _P, _I64, _P, _D,
                                          _D, _P, _P]),
    "mw_z2c": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _I64, _P]),
    "mw_c2z": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _I64, _I32, _P]),
    "mw_zreadout_check": (ctypes.c_int, [_I64, _I64, _P, _I64, _P, _I64, _P, _I64, _P, _I64,
                                         ctypes.c_double, _P, _P, _P, _P]),
}

_lib = None
_load_error = None


def load():
    """Return the loaded library, raising AcceleratorUnavailable if absent."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if _load_error is not None:
        raise AcceleratorUnavailable(_load_error)
    if not os.path.exists(LIB_PATH):
        _load_error = (f"{LIB_PATH} is not built; run `python -m momwire_amd.build` "
                       "(the port-solution path has no CPU fallback)")
        raise AcceleratorUnavailable(_load_error)
    # PyTorch-ROCm ships its own HIP runtime (soname libamdhip64.so.7, needed
    # by torch under another name).  Importing torch first makes this
    # library bind to that same runtime instead of loading a second one from
    # /opt/rocm, which would leave two runtimes competing for the device.
    import torch  # noqa: F401
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        _load_error = f"failed to load {LIB_PATH}: {e}"
        raise AcceleratorUnavailable(_load_error) from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what):
    """Map a C-ABI status code to the reference's exception types."""
    if rc == MW_OK:
        return
    msg = load().mw_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if msg else f"{what}: status {rc}"
    if rc in (MW_EINVAL,):
        raise ValueError(text)
    if rc == MW_ERANGE:
        raise ValueError(text)
    if rc == MW_EABORT:
        raise SolveAborted(text)
    if rc == MW_ESINGULAR:
        import numpy as np

        raise np.linalg.LinAlgError(text)
    raise RuntimeError(text)
