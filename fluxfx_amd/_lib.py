"""ctypes binding of libfluxhip.so (declarations in include/fluxhip.h).

The library is the only compute path: if it is missing or cannot be loaded the
import of any solver raises, there is no CPU fallback.
"""
import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfluxhip.so"

FX_OK, FX_EINVAL, FX_EDEVICE, FX_ESTATE = 0, -1, -2, -3
FX_F32, FX_BF16, FX_F16 = 0, 1, 2
FX_OP_BOX, FX_OP_FLAGS, FX_OP_EDGES, FX_OP_PACKED = 0, 1, 2, 3
FX_RESTRICT_MEAN, FX_RESTRICT_FULL = 0, 1
FX_PROLONG_TRILINEAR, FX_PROLONG_RENORM, FX_PROLONG_CONSTANT = 0, 1, 2
MAX_LEVELS = 16
FX_WS_DOUBLES = 16400  # FP64 reduction workspace (16 scalars + 16 x 1024 block partials)
FX_COARSE_CG = -1  # fx_mg_config.coarse_sweeps: coarsest level by FP64 conjugate gradients


class GPUUnavailable(RuntimeError):
    """Device failure (the reference's backend/device.py:11-12 name)."""


class fx_grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("ex", C.c_float), ("ey", C.c_float), ("ez", C.c_float)]


class fx_sources(C.Structure):
    _fields_ = [("count", C.c_int32), ("multi", C.c_int32), ("table", C.c_float * 128)]


class fx_mg_level(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("op", C.c_int32),
                ("w", C.c_float * 3), ("pad_", C.c_float), ("op_data", C.c_void_p),
                ("p", C.c_void_p), ("q", C.c_void_p), ("rhs", C.c_void_p), ("active", C.c_void_p)]


class fx_mg_config(C.Structure):
    _fields_ = [("levels", C.c_int32), ("storage", C.c_int32), ("pre", C.c_int32), ("post", C.c_int32),
                ("omega", C.c_float), ("restriction", C.c_int32), ("prolongation", C.c_int32),
                ("coarse_sweeps", C.c_int32), ("coarse_matrix", C.c_void_p), ("coarse_scale", C.c_float),
                ("separate_legs", C.c_int32), ("level", fx_mg_level * MAX_LEVELS)]


class fx_conv_desc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("c0", C.c_int32), ("c1", C.c_int32),
                ("cout", C.c_int32), ("up0", C.c_int32), ("relu", C.c_int32)]


class fx_unet(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("pad_", C.c_int32),
                ("wt", C.c_void_p * 10), ("bias", C.c_void_p * 10), ("head", C.c_void_p), ("act", C.c_void_p * 10),
                ("pool1", C.c_void_p), ("pool2", C.c_void_p), ("input", C.c_void_p)]


class fx_solve_info(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("fallback", C.c_int32),
                ("pad_", C.c_int32), ("bnorm", C.c_double), ("residual0", C.c_double),
                ("residual", C.c_double)]


class fx_solve_params(C.Structure):
    _fields_ = [("tol", C.c_float), ("max_iters", C.c_int32), ("check_from", C.c_int32), ("ref_q", C.c_int32),
                ("x_is_zero", C.c_int32), ("pad_", C.c_int32)]


class fx_velocity_args(C.Structure):
    _fields_ = [("g", fx_grid), ("dt", C.c_float), ("maccormack", C.c_int32),
                ("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p),
                ("pu", C.c_void_p), ("pv", C.c_void_p), ("pw", C.c_void_p),
                ("ou", C.c_void_p), ("ov", C.c_void_p), ("ow", C.c_void_p),
                ("density", C.c_void_p), ("temperature", C.c_void_p),
                ("thermal_lift", C.c_float), ("density_weight", C.c_float),
                ("solid", C.c_void_p), ("clearance", C.c_void_p), ("sources", C.POINTER(fx_sources)),
                ("lu", C.c_void_p), ("lv", C.c_void_p), ("lw", C.c_void_p),
                ("hu", C.c_void_p), ("hv", C.c_void_p), ("hw", C.c_void_p)]


class fx_scalar_args(C.Structure):
    _fields_ = [("g", fx_grid), ("dt", C.c_float), ("maccormack", C.c_int32), ("channels", C.c_int32),
                ("boundary", C.c_int32), ("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p),
                ("field", C.c_void_p * 3), ("pred", C.c_void_p * 3), ("out", C.c_void_p * 3),
                ("decay", C.c_float * 3), ("channel_id", C.c_int32 * 3), ("solid", C.c_void_p),
                ("clearance", C.c_void_p), ("sources", C.POINTER(fx_sources)),
                ("lo", C.c_void_p * 3), ("hi", C.c_void_p * 3)]


P = C.c_void_p
I32, I64, F32 = C.c_int32, C.c_int64, C.c_float
U32, U64, F64 = C.c_uint32, C.c_uint64, C.c_double


class fx_bricks_dev(C.Structure):
    _fields_ = [("field", C.c_void_p), ("coords", C.c_void_p), ("active", C.c_void_p), ("neighbors", C.c_void_p),
                ("side", C.c_int32), ("capacity", C.c_uint32), ("active_count", C.c_uint32),
                ("reallocations", C.c_uint32)]


class fx_projection_report(C.Structure):
    _fields_ = ([(k, C.c_uint32) for k in ("resolution", "side", "iterations", "mode", "cycles", "levels",
                                            "active_bricks", "solve_cells")] +
                [(k, C.c_uint64) for k in ("sparse_bytes", "dense_bytes", "sparse_hierarchy_bytes",
                                            "dense_hierarchy_bytes")] +
                [(k, C.c_double) for k in ("topology_ms", "sparse_gpu_ms", "dense_gpu_ms", "max_error",
                                            "pressure_error", "divergence_error", "before_rms", "after_rms",
                                            "residual_rms", "wall_error", "padding_error")])


class fx_coupled_report(C.Structure):
    _fields_ = ([("projection", fx_projection_report)] +
                [(k, C.c_uint32) for k in ("steps", "schedule", "dense_full", "global_support", "hybrid", "adaptive",
                                            "pooled", "expansions", "density_bricks", "density_capacity",
                                            "density_reallocations", "density_reuses")] +
                [(k, C.c_uint64) for k in ("density_cells", "density_bytes", "preserved_bytes", "global_copy_bytes",
                                            "peak_sparse_bytes")] +
                [(k, C.c_double) for k in ("density_error", "sparse_mass", "dense_mass", "injected_mass",
                                            "dense_after_rms", "density_topology_ms", "preflight_ms", "sparse_wall_ms",
                                            "dense_wall_ms")])


class fx_advect_report(C.Structure):
    _fields_ = ([(k, C.c_uint32) for k in ("resolution", "side", "steps", "mode", "active_bricks", "pad_")] +
                [(k, C.c_uint64) for k in ("sparse_bytes", "dense_bytes", "owned_faces", "padding_faces")] +
                [(k, C.c_double) for k in ("max_error", "rms_error", "padding_error", "topology_ms", "sparse_wall_ms",
                                            "dense_wall_ms", "sparse_mass", "dense_mass", "initial_mass")] +
                [("sparse_gpu_ms", C.c_double * 32), ("dense_gpu_ms", C.c_double * 32)])


COUPLED_OPTIONS = {"dense_full": 1, "global_support": 2, "hybrid": 4, "adaptive": 8, "pooled": 16}

SIGNATURES = {
    "fx_last_error": ([], C.c_char_p),
    "fx_projection_compare": ([U32, U32, U32, U32, U32, U64, P, P, P, P, C.POINTER(fx_projection_report), P],
                              C.c_int),
    "fx_coupled_compare": ([U32, U32, U32, U32, U32, U32, U32, U64, P, P, P, P, P, C.POINTER(fx_coupled_report), P],
                           C.c_int),
    "fx_transport_compare": ([U32, U32, U32, F32, F32, F32, U64, P, C.POINTER(fx_advect_report), P], C.c_int),
    "fx_mac_compare": ([U32, U32, U32, U32, U64, P, C.POINTER(fx_advect_report), P], C.c_int),
    "fx_version": ([], C.c_int),
    "fx_solid_raster": ([fx_grid, I32, P, P, P], C.c_int),
    "fx_build_flags": ([fx_grid, P, P, P], C.c_int),
    "fx_seed": ([fx_grid, I32, P, P, P], C.c_int),
    "fx_solid_clear": ([fx_grid, P, P, P, P], C.c_int),
    "fx_fill": ([P, I64, F32, P], C.c_int),
    "fx_scale": ([P, P, I64, F32, P], C.c_int),
    "fx_enforce_walls": ([fx_grid, P, P, P, P, P, P, P, P], C.c_int),
    "fx_divergence": ([fx_grid, P, P, P, P, P, P], C.c_int),
    "fx_jacobi_sweep": ([fx_grid, P, P, P, P, F32, F32, P], C.c_int),
    "fx_gradient": ([fx_grid, P, P, P, P, P, P, P, P, F32, P], C.c_int),
    "fx_gradient_split": ([fx_grid, P, P, P, P, P, P, P, P, P, P, F32, P], C.c_int),
    "fx_increment_rhs": ([fx_grid, P, P, P, P, P, P], C.c_int),
    "fx_conv3d_forward": ([C.POINTER(fx_conv_desc), P, P, P, P, P, P], C.c_int),
    "fx_conv3d_weight_grad": ([C.POINTER(fx_conv_desc), P, P, P, P, P, I64, P, P, I32, P], C.c_int),
    "fx_conv3d_weight_grad_workspace": ([C.POINTER(fx_conv_desc)], C.c_int64),
    "fx_conv_pack_weights": ([I32, I32, I32, I32, P, P, P], C.c_int),
    "fx_relu_mask": ([I64, P, P, P], C.c_int),
    "fx_pool": ([I32, I32, I32, I32, P, P, P], C.c_int),
    "fx_pool_backward_add": ([I32, I32, I32, I32, P, P, P, P], C.c_int),
    "fx_concat_backward": ([I32, I32, I32, I32, I32, P, P, P, P], C.c_int),
    "fx_unet_input": ([I64, P, P, P, P, P], C.c_int),
    "fx_unet_head": ([I64, P, P, P, P, P, F32, P, P], C.c_int),
    "fx_unet_head_backward": ([I64, P, P, P, P, P, P], C.c_int),
    "fx_masked_mse": ([I64, P, P, P, F32, P, P, P], C.c_int),
    "fx_adam": ([I64, P, P, P, P, I32, F32, F32, F32, F32, P], C.c_int),
    "fx_to_bf16": ([I64, P, P, P], C.c_int),
    "fx_unet_forward": ([C.POINTER(fx_unet), P, P, P, F32, I32, P, P], C.c_int),
    "fx_learned_pcg_solve": ([C.POINTER(fx_unet), C.POINTER(fx_mg_config), P, P, P, P, P, P, P, P, P, P, P,
                              F32, I32, I32, C.POINTER(fx_solve_params), P, C.POINTER(fx_solve_info), P], C.c_int),
    "fx_combine": ([P, P, P, I64, P], C.c_int),
    "fx_curl": ([fx_grid, P, P, P, P, P], C.c_int),
    "fx_enforce_walls_moving": ([fx_grid, P, P, P, P, P, P, P, P, P, P, P], C.c_int),
    "fx_divergence_moving": ([fx_grid, P, P, P, P, P, P, P, P, P], C.c_int),
    "fx_gradient_moving": ([fx_grid, P, P, P, P, P, P, P, P, P, P, P, F32, P], C.c_int),
    "fx_gradient_split_moving": ([fx_grid, P, P, P, P, P, P, P, P, P, P, P, P, P, F32, P], C.c_int),
    "fx_owner_raster": ([fx_grid, I32, P, I32, P, P], C.c_int),
    "fx_wall_velocity": ([fx_grid, P, I32, P, P, P, P, P], C.c_int),
    "fx_remap_scalar": ([fx_grid, P, P, P, P, P], C.c_int),
    "fx_remap_faces": ([fx_grid, P, P, P, P, P, P, P, P, P, P, P], C.c_int),
    "fx_moving_update_box": ([fx_grid, P, I32, P, P, P, P, P, P, P, P, P, P, P, P], C.c_int),
    "fx_remap_box": ([fx_grid, P, P, P, P, P, P, P, P, P, P, P, P, P], C.c_int),
    "fx_confine": ([fx_grid, P, P, P, P, P, P, P, F32, F32, F32, P], C.c_int),
    "fx_confine_split": ([fx_grid, P, P, P, P, P, P, P, P, F32, F32, F32, P], C.c_int),
    "fx_bricks_create": ([I32, U32, U64, C.POINTER(P)], C.c_int),
    "fx_bricks_destroy": ([P], C.c_int),
    "fx_bricks_close": ([P], C.c_int),
    "fx_bricks_activate": ([P, I32, I32, I32, C.POINTER(U32)], C.c_int),
    "fx_bricks_deactivate": ([P, I32, I32, I32, C.POINTER(I32)], C.c_int),
    "fx_bricks_lookup": ([P, I32, I32, I32, C.POINTER(I32)], C.c_int),
    "fx_bricks_neighbors": ([P, I32, I32, I32, C.POINTER(I32)], C.c_int),
    "fx_bricks_activate_box": ([P, I32, I32, I32, I32, I32, I32], C.c_int),
    "fx_bricks_activate_flags": ([P, P, I32, I32, I32], C.c_int),
    "fx_bricks_update_regions": ([P, P, I32, I32, F64, I32, I32], C.c_int),
    "fx_bricks_count": ([P, C.POINTER(U32), C.POINTER(U32)], C.c_int),
    "fx_bricks_snapshot": ([P, P, U32, C.POINTER(U32)], C.c_int),
    "fx_bricks_host_bytes": ([P, C.POINTER(U64)], C.c_int),
    "fx_bricks_reserve": ([P, U32, U32, C.POINTER(U32)], C.c_int),
    "fx_bricks_sync": ([P, P], C.c_int),
    "fx_bricks_device": ([P, C.POINTER(fx_bricks_dev)], C.c_int),
    "fx_bricks_gather": ([P, I32, I32, I32, P, P], C.c_int),
    "fx_bricks_scatter": ([P, I32, I32, I32, P, P], C.c_int),
    "fx_brick_support": ([I32, I32, I32, I32, P, F32, P, P], C.c_int),
    "fx_bricks_probe": ([P, C.POINTER(U32), C.POINTER(F32), P], C.c_int),
    "fx_brick_required": ([I32, I32, P, P, P, P, P, P, P], C.c_int),
    "fx_bricks_transport": ([P, I32, P, P, P, P], C.c_int),
    "fx_mg_vcycle": ([C.POINTER(fx_mg_config), F32, P], C.c_int),
    "fx_mg_smooth": ([C.POINTER(fx_mg_config), I32, I32, F32, P], C.c_int),
    "fx_mg_residual": ([C.POINTER(fx_mg_config), I32, F32, P, P], C.c_int),
    "fx_mg_residual_restrict": ([C.POINTER(fx_mg_config), I32, F32, P], C.c_int),
    "fx_mg_restrict": ([C.POINTER(fx_mg_config), I32, P, P, P], C.c_int),
    "fx_mg_prolong": ([C.POINTER(fx_mg_config), I32, P, P, P, P], C.c_int),
    "fx_mg_direct": ([C.POINTER(fx_mg_config), P], C.c_int),
    "fx_mg_coarse_cg": ([C.POINTER(fx_mg_config), P, P], C.c_int),
    "fx_mg_rebuild_masked": ([C.POINTER(fx_mg_config), F64, F64, F64, P, P], C.c_int),
    "fx_project_mean": ([fx_grid, P, P, P, P], C.c_int),
    "fx_project_mean_noalloc": ([fx_grid, P, P, P], C.c_int),
    "fx_residual_norm": ([fx_grid, P, P, P, C.POINTER(C.c_double), C.POINTER(C.c_double), P, P], C.c_int),
    "fx_mg_solve": ([C.POINTER(fx_mg_config), C.POINTER(fx_solve_params), P, C.POINTER(fx_solve_info), P],
                    C.c_int),
    "fx_mixed_solve": ([C.POINTER(fx_mg_config), P, P, P, C.POINTER(fx_solve_params), P,
                        C.POINTER(fx_solve_info), C.POINTER(C.POINTER(C.c_float)), P], C.c_int),
    "fx_pcg_solve": ([C.POINTER(fx_mg_config), P, P, P, P, P, P, C.POINTER(fx_solve_params), P,
                      C.POINTER(fx_solve_info), P], C.c_int),
    "fx_transport_fast_paths": ([C.c_int32], C.c_int),
    "fx_face_predict": ([fx_grid, F32, P, P, P, P, P, P, P, P, P], C.c_int),
    "fx_face_predict_bounds": ([fx_grid, F32, P, P, P, P, P, P, P, P, P], C.c_int),
    "fx_build_clearance": ([fx_grid, P, P, P, P], C.c_int),
    "fx_velocity_advect": ([C.POINTER(fx_velocity_args), P], C.c_int),
    "fx_scalar_predict": ([fx_grid, F32, P, P, P, I32, C.POINTER(P), C.POINTER(P), P, P, P], C.c_int),
    "fx_scalar_predict_bounds": ([fx_grid, F32, P, P, P, I32, C.POINTER(P), C.POINTER(P), C.POINTER(P),
                                  C.POINTER(P), P, P, P], C.c_int),
    "fx_scalar_advect": ([C.POINTER(fx_scalar_args), P], C.c_int),
    "fx_combustion": ([I64, F32, F32, F32, F32, F32, P, P, P, P, P, P, P, P], C.c_int),
    "fx_max_abs": ([I32, C.POINTER(P), C.POINTER(I64), P, P], C.c_int),
    "fx_active_scan": ([fx_grid, P, P, P], C.c_int),
    "fx_divergence_bound": ([fx_grid, P, P, P, P, P, P, P, P, P, P, P], C.c_int),
    "fx_active_plan": ([fx_grid, F32, P, I32, I32, C.POINTER(P), I32, C.POINTER(P), C.POINTER(fx_sources), P, P, P],
                       C.c_int),
    "fx_scalar_predict_bounds_active": ([fx_grid, F32, P, P, P, I32, C.POINTER(P), C.POINTER(P), C.POINTER(P),
                                         C.POINTER(P), P, P, P, P], C.c_int),
    "fx_scalar_advect_active": ([C.POINTER(fx_scalar_args), P, C.POINTER(P), P], C.c_int),
    "fx_combustion_active": ([fx_grid, F32, F32, F32, F32, F32, P, P, P, P, P, P, P, P, C.POINTER(P), P], C.c_int),
}

_lib = None


def load(path=None):
    """Load libfluxhip.so once and attach the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(path or LIB_PATH)
    if not path.exists():
        raise GPUUnavailable(f"{path} is not built; run fluxfx_amd.build.build() first")
    lib = C.CDLL(str(path))
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def exported_names():
    return list(SIGNATURES)


def check(code):
    """Map a C-ABI status to the reference's exception types."""
    if code == FX_OK:
        return
    msg = _lib.fx_last_error().decode(errors="replace") if _lib is not None else "libfluxhip error"
    if code == FX_EINVAL:
        raise ValueError(msg)
    if code == FX_ESTATE:
        raise RuntimeError(msg)
    raise GPUUnavailable(msg)


def call(name, *args):
    lib = load()
    check(getattr(lib, name)(*args))


def make_grid(shape, extent):
    return fx_grid(int(shape[0]), int(shape[1]), int(shape[2]),
                   float(extent[0]), float(extent[1]), float(extent[2]))


def make_sources(rows, multi):
    s = fx_sources()
    s.count = len(rows)
    s.multi = 1 if multi else 0
    flat = [float(v) for row in rows for v in row]
    if len(rows) > 8:
        raise ValueError("At most eight sources are supported")
    for i, v in enumerate(flat):
        s.table[i] = v
    return s
