"""ctypes binding of libncx.so (the C ABI declared in include/ncx.h)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libncx.so")

_lib = None

P = C.c_void_p
SZ = C.c_size_t
I32 = C.c_int32
U32 = C.c_uint32
F32 = C.c_float
PP6 = P * 6

# Every symbol the header declares, with (restype, argtypes).
SIGNATURES = {
    "ncx_last_error": (C.c_char_p, []),
    "ncx_init": (C.c_int, [P, C.c_int, P]),
    "ncx_free": (C.c_int, [P]),
    "ncx_set_stream": (C.c_int, [P, P]),
    "ncx_synchronize": (C.c_int, [P]),
    "ncx_bp_create": (C.c_int, [P, F32, P]),
    "ncx_bp_destroy": (C.c_int, [P]),
    "ncx_bp_create_proxies": (C.c_int, [P, P, SZ, P]),
    "ncx_bp_set_aabbs_dev": (C.c_int, [P, P, P, SZ]),
    "ncx_bp_set_aabbs": (C.c_int, [P, P, P, SZ]),
    "ncx_bp_remove": (C.c_int, [P, P, SZ]),
    "ncx_bp_update": (C.c_int, [P]),
    "ncx_bp_num_pairs": (C.c_int, [P, P]),
    "ncx_bp_pairs": (C.c_int, [P, P, SZ, P]),
    "ncx_bp_pairs_dev": (C.c_int, [P, P, P, P]),
    "ncx_bp_copy_pairs_dev": (C.c_int, [P, P, P, SZ, P]),
    "ncx_bp_events": (C.c_int, [P, P, SZ, P, P, SZ, P]),
    "ncx_bp_proxy_aabb": (C.c_int, [P, U32, P]),
    "ncx_shapes_create": (C.c_int, [P, P, SZ, P, P, P, SZ, P]),
    "ncx_shapes_destroy": (C.c_int, [P]),
    "ncx_shapes_count": (C.c_int, [P, P]),
    "ncx_shapes_tuning": (C.c_int, [P, C.c_int, P, P]),
    "ncx_shapes_set_tuning": (C.c_int, [P, C.c_int, P]),
    "ncx_shapes_aabbs_dev": (C.c_int, [P, P, F32, P]),
    "ncx_contact_batch": (C.c_int, [P, P, P, P, SZ, F32, P, P, P, P, P, P, P]),
    "ncx_contact_batch_capped": (C.c_int, [P, P, P, P, SZ, F32, P, I32, P, P, P, P, P, P]),
    "ncx_distance_batch": (C.c_int, [P, P, P, P, SZ, P, P]),
    "ncx_distance_batch_capped": (C.c_int, [P, P, P, P, SZ, P, C.c_int32, P, P]),
    "ncx_time_of_impact_batch": (C.c_int, [P, P, P, P, P, SZ, F32, F32, P, P, P, P, P, P, P]),
    "ncx_shapes_cast_rays": (C.c_int, [P, P, P, P, SZ, F32, C.c_int, P, P, P]),
    "ncx_mesh_contact_batch": (C.c_int, [P, P, P, P, P, SZ, F32, P, P, P, P, P, P]),
    "ncx_mesh_distance_batch": (C.c_int, [P, P, P, P, P, SZ, P]),
    "ncx_hull_triangulate": (C.c_int, [P, SZ, P, SZ, P]),
    "ncx_hull_polyhedra_check": (C.c_int, [P, P, P, SZ, P, C.c_int]),
    "ncx_hull_polyhedron": (C.c_int, [P, SZ, P, SZ, P, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "ncx_manifolds_create": (C.c_int, [P, SZ, P]),
    "ncx_manifolds_destroy": (C.c_int, [P]),
    "ncx_manifolds_reset": (C.c_int, [P, P, SZ]),
    "ncx_manifolds_dirs": (C.c_int, [P, P, SZ, P]),
    "ncx_contact_manifolds": (C.c_int, [P, P, P, P, P, P, SZ, F32, F32, F32, P, P, P, P, P, P, P]),
    "ncx_trimesh_create": (C.c_int, [P, P, SZ, P, SZ, P]),
    "ncx_trimesh_destroy": (C.c_int, [P]),
    "ncx_trimesh_cast_rays": (C.c_int, [P, P, P, SZ, F32, P, P, P]),
    "ncx_trimesh_cast_stats": (C.c_int, [P, P, P, SZ, F32, P, P]),
    "ncx_trimesh_cast_rays_async": (C.c_int, [P, P, P, SZ, F32, P, P, P]),
    "ncx_trimesh_check": (C.c_int, [P]),
    "ncx_trimesh_cast_rays_noalloc": (C.c_int, [P, P, P, SZ, F32, P, P, P]),
    "ncx_trimesh_set_timing": (C.c_int, [P, C.c_int]),
    "ncx_trimesh_kernel_ms": (C.c_int, [P, P]),
    "ncx_convex_hull": (C.c_int, [P, SZ, P, P]),
    "ncx_convex_hull_batch": (C.c_int, [P, P, P, SZ, P, P, C.c_int]),
    "ncx_mlp_create": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_uint64, P]),
    "ncx_mlp_destroy": (C.c_int, [P]),
    "ncx_mlp_num_params": (C.c_int, [P, P]),
    "ncx_mlp_get_params": (C.c_int, [P, P]),
    "ncx_mlp_set_params": (C.c_int, [P, P]),
    "ncx_mlp_forward": (C.c_int, [P, P, SZ, P]),
    "ncx_mlp_train_step": (C.c_int, [P, P, P, SZ, F32, P]),
    "ncx_mlp_grad": (C.c_int, [P, P, P, SZ, P, P]),
    "ncx_mlp_predict_dirs": (C.c_int, [P, P, P, P, P, SZ, P]),
    "ncx_features": (C.c_int, [P, P, P, P, P, SZ, P]),
    "ncx_shape_descriptors": (C.c_int, [P, P]),
    "ncx_world_create": (C.c_int, [P, P, F32, F32, SZ, P]),
    "ncx_world_destroy": (C.c_int, [P]),
    "ncx_world_set_warmstart": (C.c_int, [P, P]),
    "ncx_world_set_rebuild_every": (C.c_int, [P, C.c_int]),
    "ncx_world_step": (C.c_int, [P, P, C.c_int]),
    "ncx_world_set_contacts": (C.c_int, [P, C.c_int]),
    "ncx_world_query_stats": (C.c_int, [P, P]),
    "ncx_world_spill_stats": (C.c_int, [P, P]),
    "ncx_world_sync": (C.c_int, [P, P, P]),
    "ncx_world_outputs": (C.c_int, [P, P, P, P, P, P, P, P, P]),
    "ncx_world_stored_aabbs": (C.c_int, [P, P]),
    "ncx_world_warmstart_dirs": (C.c_int, [P, P]),
    "ncx_world_set_temporal_warmstart": (C.c_int, [P, C.c_int]),
    "ncx_world_temporal_dirs": (C.c_int, [P, P, P]),
    "ncx_world_interferences_with_rays": (C.c_int, [P, P, P, SZ, F32, C.c_int, P, P, P, P, SZ, P]),
    "ncx_world_first_interference_with_rays": (C.c_int, [P, P, P, SZ, F32, P, P, P]),
    "ncx_world_interferences_with_aabbs": (C.c_int, [P, P, SZ, P, P, SZ, P]),
    "ncx_bp_remove_with_pairs": (C.c_int, [P, P, SZ, P, SZ, P]),
    "ncx_bp_set_collision_groups": (C.c_int, [P, P, P, P, P, SZ]),
    "ncx_bp_interferences_with_aabbs": (C.c_int, [P, P, SZ, P, P, SZ, P]),
    "ncx_bp_interferences_with_points": (C.c_int, [P, P, SZ, P, P, SZ, P]),
    "ncx_bp_interferences_with_rays": (C.c_int, [P, P, SZ, F32, P, P, SZ, P]),
}


class NcxError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Loads libncx.so and declares its signatures.  Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NcxError(f"{path} is missing: build it with `python -m ncollide_amd.build` "
                       "or __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().ncx_last_error()
        raise NcxError(f"libncx error {rc}: {msg.decode() if msg else ''}")
    return rc


def ptr(t):
    """Raw pointer of a tensor / ndarray (None stays NULL)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return C.c_void_p(t.data_ptr())
    return t.ctypes.data_as(C.c_void_p)


def ptr6(arrs):
    return PP6(*[C.c_void_p(a.data_ptr()) if hasattr(a, "data_ptr") else a.ctypes.data_as(C.c_void_p)
                 for a in arrs])


def ptrs(arrs):
    """Array of raw pointers (device tensors or host arrays) for `const float* const x[k]` arguments."""
    return (P * len(arrs))(*[C.c_void_p(a.data_ptr()) if hasattr(a, "data_ptr") else a.ctypes.data_as(C.c_void_p)
                             for a in arrs])


def soa_columns(rows, k, device):
    """(n, k) host or device array -> k contiguous float32 device columns."""
    import torch
    t = torch.as_tensor(rows, device=device, dtype=torch.float32).reshape(-1, k)
    return [t[:, d].contiguous() for d in range(k)]
