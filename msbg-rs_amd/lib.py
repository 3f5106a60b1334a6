"""ctypes binding of the C ABI in include/msbg_hip.h (libmsbg_hip.so).

`load()` returns a `HipLib` whose methods raise `MsbgError` on a non-zero
status.  The library must have been built (msbg-rs_amd/build.py); there is no
fallback -- a missing library is an error.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmsbg_hip.so")

vp, i64, i32, f32, u64 = C.c_void_p, C.c_int64, C.c_int, C.c_float, C.c_uint64
PP = C.POINTER(C.c_void_p)

# name -> argtypes (all return int status unless listed in _RESTYPE)
_SIGS = {
    "msbg_last_error": [],
    "msbg_device_count": [C.POINTER(C.c_int)],
    "msbg_set_stream": [vp],
    "msbg_synchronize": [],
    "msbg_malloc": [PP, C.c_size_t],
    "msbg_free": [vp],
    "msbg_memcpy_htod": [vp, vp, C.c_size_t],
    "msbg_memcpy_dtoh": [vp, vp, C.c_size_t],
    "msbg_memcpy_dtod": [vp, vp, C.c_size_t],
    "msbg_fill_u32": [vp, i64, C.c_uint32],
    "msbg_event_create": [PP],
    "msbg_event_destroy": [vp],
    "msbg_event_record": [vp],
    "msbg_event_elapsed_ms": [vp, vp, C.POINTER(C.c_float)],
    "msbg_grid_new": [i64, i64, i64, i32, f32, f32, i64, PP],
    "msbg_grid_free": [vp],
    "msbg_grid_dims": [vp, vp],
    "msbg_grid_ensure_blocks": [vp, vp, i64],
    "msbg_grid_fill_blocks": [vp, vp, i64, f32],
    "msbg_grid_set_full_block": [vp, i64],
    "msbg_grid_set_empty_block": [vp, i64],
    "msbg_grid_set_empty_value": [vp, f32],
    "msbg_grid_set_full_value": [vp, f32],
    "msbg_grid_write_blocks": [vp, vp, i64, vp],
    "msbg_grid_read_blocks": [vp, vp, i64, vp],
    "msbg_grid_load": [vp, vp, vp, i64],
    "msbg_grid_store": [vp, vp, vp, i64, vp],
    "msbg_grid_block_slots": [vp, vp],
    "msbg_grid_copy": [vp, vp],
    "msbg_grid_value_block_ptr": [vp, i64, PP],
    "msbg_grid_device_views": [vp, PP, PP],
    "msbg_color_buckets": [vp, vp, i64, vp, vp],
    "msbg_sweep": [vp, vp, i64, i32, f32, i32, i32],
    "msbg_sweep_plan_new": [vp, vp, i64, PP],
    "msbg_sweep_plan_free": [vp],
    "msbg_sweep_density": [vp, vp, vp, i64, i32, f32, i32, i32],
    "msbg_sweep_plan_new_density": [vp, vp, vp, i64, PP],
    "msbg_sweep_density8": [vp, vp, vp, i64, i32, f32, i32, i32, C.c_uint32],
    "msbg_sweep_plan_new_density8": [vp, vp, vp, i64, C.c_uint32, PP],
    "msbg_sweep_plan_run_async": [vp, i32, f32, i32, i32],
    "msbg_sweep_plan_run_timed": [vp, i32, f32, i32, i32, C.POINTER(C.c_float), C.POINTER(C.c_int64)],
    "msbg_sweep_plan_run_graph_async": [vp, i32, f32, i32, i32],
    "msbg_halo_topology": [vp, vp, vp, vp, i64, i32, vp],
    "msbg_halo_gather": [vp, vp, i64, i32, i32, i32, vp],
    "msbg_halo_gather_multires": [vp, vp, vp, vp, i64, i32, i32, i32, vp],
    "msbg_sweep_multires": [vp, vp, vp, vp, i64, i32, f32, i32, i32],
    "msbg_sweep_plan_new_multires": [vp, vp, vp, vp, i64, PP],
    "msbg_sweep_levels": [vp, i32, vp, vp, i64, i32, f32, i32, i32],
    "msbg_sweep_plan_new_levels": [vp, i32, vp, vp, i64, PP],
    "msbg_downsample_avg": [vp, vp, vp, i64],
    "msbg_regularize_levels": [vp, i64, i64, i64, i32],
    "msbg_morton_keys": [vp, i64, vp],
    "msbg_radix_sort_pairs": [vp, vp, i64, i32],
    "msbg_morton_sort": [vp, i64, vp, vp],
    "msbg_particle_bids": [i64, i64, i64, vp, i64, vp],
    "msbg_bucket_indices": [vp, i64, i64, vp, vp, vp, C.POINTER(C.c_int64)],
    "msbg_active_blocks": [i64, i64, i64, vp, i64, f32, vp, C.POINTER(C.c_int64)],
    "msbg_active_blocks_noalloc": [i64, i64, i64, vp, i64, f32, vp, C.POINTER(C.c_int64)],
    "msbg_gather_rows": [vp, vp, i64, i64, vp],
    "msbg_morton_sort_particles": [vp, vp, vp, i64, vp, vp, vp, vp, vp],
    "msbg_sample": [vp, vp, i64, i32, i32, i32, vp],
    "msbg_mac_sample": [vp, vp, i64, i32, i32, vp],
    "msbg_g2p_flip": [vp, vp, vp, vp, i64, i32, i32, f32, vp],
    "msbg_g2p_blend": [vp, vp, vp, i64, f32, vp],
    "msbg_particles_to_grid_mac": [vp, vp, vp, vp, vp, i64, f32, C.POINTER(C.c_int64)],
    "msbg_p2g_bspline": [i32, vp, vp, vp, vp, vp, i64, C.POINTER(C.c_int64)],
    "msbg_g2p_bspline": [i32, vp, vp, i64, vp],
    "msbg_density_splat": [vp, vp, i64, vp, i64, f32, C.POINTER(C.c_int64)],
    "msbg_surface_place": [vp, i64, vp, f32, i64, i64, i64, i64, f32, vp, vp, i64, C.POINTER(C.c_int64)],
    "msbg_density_finalize": [vp, vp, i64, vp, i64, f32, f32],
    "msbg_block_topology": [vp, i64, i64, i64, i32, vp, vp, vp, vp, C.POINTER(C.c_int64)],
    "msbg_mac_divergence": [vp, vp, vp, i64],
    "msbg_mac_subtract_gradient": [vp, vp, vp, vp, i64, f32],
    "msbg_mac_subtract_gradient_pair": [vp, vp, vp, vp, vp, i64, f32],
    "msbg_poisson_new": [vp, vp, i64, PP],
    "msbg_poisson_free": [vp],
    "msbg_poisson_info": [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "msbg_poisson_jacobi": [vp, i32, vp, vp, vp, f32],
    "msbg_poisson_rbgs": [vp, i32, vp, vp, i32, f32],
    "msbg_poisson_jacobi_sweeps": [vp, i32, vp, vp, vp, f32, i32],
    "msbg_poisson_residual": [vp, i32, vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    "msbg_poisson_residual_pair": [vp, vp, vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    "msbg_poisson_convert": [i32, vp, i32, vp, i64],
    "msbg_poisson_solve": [vp, vp, vp, vp, vp],
    "msbg_learned_param_count": [C.POINTER(C.c_int)],
    "msbg_learned_new": [PP],
    "msbg_learned_free": [vp],
    "msbg_learned_set_params": [vp, vp, f32],
    "msbg_learned_get_params": [vp, vp, C.POINTER(C.c_float)],
    "msbg_learned_forward": [vp, vp, i64, vp],
    "msbg_learned_loss_grad": [vp, vp, vp, i64, C.POINTER(C.c_double), vp],
    "msbg_learned_train": [vp, vp, vp, i64, i32, i32, f32, u64, vp],
    "msbg_poisson_net_input": [vp, vp, vp, f32, vp],
    "msbg_poisson_error_target": [vp, vp, vp, f32, vp],
    "msbg_poisson_learned_correct": [vp, vp, vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int)],
}
_RESTYPE = {"msbg_last_error": C.c_char_p}


class PoissonParams(C.Structure):
    """msbg_poisson_params (include/msbg_hip.h)."""
    _fields_ = [("storage", C.c_int), ("refine", C.c_int), ("omega", C.c_float), ("sweeps_per_check", C.c_int),
                ("max_sweeps", C.c_int), ("flags", C.c_int), ("tol", C.c_double), ("learned", C.c_void_p),
                ("phi", C.c_void_p), ("x_lo", C.c_void_p)]


class PoissonStats(C.Structure):
    """msbg_poisson_stats (include/msbg_hip.h)."""
    _fields_ = [("sweeps", C.c_int), ("checks", C.c_int), ("converged", C.c_int), ("corrections", C.c_int),
                ("accepted", C.c_int), ("levels", C.c_int), ("rel_residual", C.c_double), ("r2", C.c_double),
                ("b2", C.c_double), ("ms", C.c_float)]


class MsbgError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__("%s failed (%d): %s" % (fn, code, msg))
        self.code = code


def ptr(a):
    """Address of a numpy array (host) or an int device pointer; None -> NULL."""
    if a is None:
        return None
    if isinstance(a, (int, np.integer)):
        return int(a)
    if isinstance(a, DeviceArray):
        return a.ptr
    return a.ctypes.data


class HipLib:
    def __init__(self, path=LIB_PATH):
        self.path = path
        self.cdll = C.CDLL(path)
        for name, args in _SIGS.items():
            fn = getattr(self.cdll, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)

    def symbols(self):
        return list(_SIGS)

    def call(self, name, *args):
        rc = getattr(self.cdll, name)(*args)
        if rc != 0:
            raise MsbgError(name, rc, self.cdll.msbg_last_error().decode())
        return rc

    def device_count(self):
        n = C.c_int(0)
        self.call("msbg_device_count", C.byref(n))
        return n.value

    def synchronize(self):
        self.call("msbg_synchronize")

    # -- device memory ---------------------------------------------------------
    def malloc(self, nbytes):
        p = C.c_void_p()
        self.call("msbg_malloc", C.byref(p), nbytes)
        return p.value or 0

    def free(self, p):
        self.call("msbg_free", p)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        if a.nbytes:
            self.call("msbg_memcpy_htod", d.ptr, a.ctypes.data, a.nbytes)
        return d

    def empty(self, shape, dtype):
        return DeviceArray(self, shape, dtype)

    def event(self):
        p = C.c_void_p()
        self.call("msbg_event_create", C.byref(p))
        return p.value

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        self.call("msbg_event_elapsed_ms", a, b, C.byref(ms))
        return ms.value


class DeviceArray:
    """A device buffer owned by the library allocator (msbg_malloc)."""

    def __init__(self, lib, shape, dtype):
        self.lib = lib
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = lib.malloc(self.nbytes) if self.nbytes else 0

    def get(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            self.lib.call("msbg_memcpy_dtoh", out.ctypes.data, self.ptr, self.nbytes)
        return out

    def set(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        if self.nbytes:
            self.lib.call("msbg_memcpy_htod", self.ptr, a.ctypes.data, self.nbytes)

    def free(self):
        if self.ptr:
            self.lib.free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_LIB = None


def load(path=LIB_PATH):
    """Load libmsbg_hip.so; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise RuntimeError("libmsbg_hip.so is not built: run __graft_entry__.build() first")
        _LIB = HipLib(path)
    return _LIB
