"""ctypes bindings of the CPU oracle (oracle/liborc.so) -- test infrastructure."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
i64, i32, f32 = C.c_int64, C.c_int, C.c_float
f32p_or_null = C.c_void_p


def _sig(lib, name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(ROOT, "oracle", "liborc.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle/liborc.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    _sig(lib, "orc_morton3", C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64)
    _sig(lib, "orc_color_buckets", None, i64, i64, i64, i32, i64p, i64, i64p, i64p)
    _sig(lib, "orc_halo_fill", None, i64, i64, i64, i32, f32p, i32p, i64, i32, i32, i32, f32p)
    _sig(lib, "orc_halo_fill_multires", None, i64, i64, i64, i32, f32p, i32p, f32p, i32p, u8p, i64, i32, i32,
         i32, f32p)
    for k in ("laplacian", "meancurv", "bilaplacian"):
        _sig(lib, "orc_kernel_" + k, None, i32, f32p, f32, i32, f32p)
    _sig(lib, "orc_sweep", None, i64, i64, i64, f32p, i32p, i64p, i64, i32, f32, i32, i32, i32, i32)
    _sig(lib, "orc_set_multires_ftz", None, i32)
    _sig(lib, "orc_sweep_multires", None, i64, i64, i64, i32, f32p, i32p, f32p, i32p, u8p, i64p, i64, i32, f32,
         i32, i32, i32)
    _sig(lib, "orc_halo_fill_level", None, i64p, C.c_void_p, C.c_void_p, u8p, i32, i32, i64, i32, f32p)
    _sig(lib, "orc_sweep_levels", None, i64p, C.c_void_p, C.c_void_p, u8p, i32, i64p, i64, i32, f32, i32, i32, i32)
    _sig(lib, "orc_halo_topology", None, i64, i64, i64, i32, i32p, C.c_void_p, C.c_void_p, i64, i32, i32p)
    _sig(lib, "orc_downsample_avg", None, i32, f32p, i32p, f32p, i32p, i64p, i64)
    _sig(lib, "orc_sample_coarse", f32, i64, i64, i64, i32, f32p, i32p, i32, i32, i32)
    _sig(lib, "orc_regularize", None, u8p, i64, i64, i64, i32)
    _sig(lib, "orc_active_blocks", i64, i64, i64, i64, f32p, i64, f32, i64p)
    _sig(lib, "orc_bucket_indices", i64, i64p, i64, i64, i64p, i64p, i64p)
    _sig(lib, "orc_particle_bids", None, i64, i64, i64, f32p, i64, i64p)
    _sig(lib, "orc_morton_keys", None, f32p, i64, u32p)
    _sig(lib, "orc_stable_sort_perm", None, u32p, i64, u32p)
    _sig(lib, "orc_sample", None, i64, i64, i64, f32p, i32p, f32p, i64, i32, i32, i32, f32p_or_null,
         f32p_or_null, f32p_or_null)
    _sig(lib, "orc_mac_sample", None, i64, i64, i64, f32p, i64, i32p, f32p, i64, i32, i32, f32p)
    _sig(lib, "orc_g2p_blend", None, f32p, f32p, f32p, i64, f32, f32p)
    _sig(lib, "orc_p2g_mac", None, i64, i64, i64, f32p, f32p, f32p, i64p, i64p, i64, f32, f32p, i64, i32p,
         i64p, i64, i32)
    _sig(lib, "orc_p2g_mac_naive", None, i64, i64, i64, f32p, f32p, f32p, i64, f32, f32p)
    _sig(lib, "orc_p2g_mac_f64", None, i64, i64, i64, f32p, f32p, f32p, i64p, i64p, i64, f32,
         np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS"), i64, i32p, i64p, i64)
    _sig(lib, "orc_bspline_weight", f32, i32, f32, i64)
    _sig(lib, "orc_p2g_bspline", None, i32, i64, i64, i64, f32p, f32p, C.c_void_p, i64, i32p, i64, i64p, i64, f32p)
    _sig(lib, "orc_g2p_bspline", None, i32, i64, i64, i64, f32p, i64, i32p, f32p, i64, f32p)
    u16 = C.c_uint16
    _sig(lib, "orc_f32_to_f16", u16, f32)
    _sig(lib, "orc_f16_to_f32", f32, u16)
    _sig(lib, "orc_f32_to_bf16", u16, f32)
    _sig(lib, "orc_bf16_to_f32", f32, u16)
    vp = C.c_void_p
    _sig(lib, "orc_poisson_residual", None, i64, i64, i64, i32, vp, vp, i32p, i64p, i64, vp,
         C.POINTER(C.c_double), C.POINTER(C.c_double))
    _sig(lib, "orc_jacobi", None, i64, i64, i64, i32, vp, vp, vp, i32p, i64p, i64, f32)
    _sig(lib, "orc_rbgs", None, i64, i64, i64, i32, vp, vp, i32p, i64p, i64, i32, f32)
    _sig(lib, "orc_convert", None, i32, vp, i32, vp, i64)
    _sig(lib, "orc_poisson_solve", None, i64, i64, i64, f32p, f32p, i64, i32p, i64p, i64, i32, i32, f32, i32, i32,
         C.c_double, np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS"))
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    _sig(lib, "orc_cg64", i32, i64, i64, i64, f64p, f64p, i64, i32p, i64p, i64, C.c_double, i32)
    _sig(lib, "orc_mac_divergence", None, i64, i64, i64, f32p, i32p, f32p, i32p, f32p, i32p, f32p, i32p, i64p, i64,
         i32)
    _sig(lib, "orc_sweep_density", None, i64, i64, i64, np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS"), i64,
         i32p, i64p, i64, i32, f32, i32, i32)
    _sig(lib, "orc_density8_uniform", f32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32)
    _sig(lib, "orc_density8_code", C.c_uint8, f32, f32)
    _sig(lib, "orc_sweep_density8", None, i64, i64, i64, u8p, i64, i32p, i64p, i64, i32, f32, i32, i32, C.c_uint32,
         C.c_uint32)
    _sig(lib, "orc_net_param_count", i32)
    _sig(lib, "orc_net_forward", None, f32p, f32p, i64, i32, f32p)
    _sig(lib, "orc_net_loss_grad", C.c_double, f32p, f32p, f32p, i64, i32, f64p)
    _sig(lib, "orc_net_input", None, i64, i64, i64, i32p, i64p, i64, f32p, f32, f32p, i32p,
         np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS"))
    u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
    _sig(lib, "orc_domain_bounds", None, i64, i64, i64, f32p)
    _sig(lib, "orc_surface_place", i64, f32p, i64, f32p, f32, i64, i64, i64, i64, f32, f32p, i64p, i64)
    _sig(lib, "orc_density_finalize", None, i32p, u16p, i64p, i64, f32, f32)
    _sig(lib, "orc_block_topology", i64, i64, i64, i64, u8p, i32, vp, u16p, i64p, i64p)
    _sig(lib, "orc_mac_subtract_gradient", None, i64, i64, i64, f32p, i32p, f32p, i32p, f32p, i32p, f32p, i32p, i64p,
         i64, f32)
    _sig(lib, "orc_mac_subtract_gradient_pair", None, i64, i64, i64, f32p, i32p, f32p, i32p, f32p, i32p, f32p, f32p,
         i32p, i64p, i64, f32)
    _sig(lib, "orc_poisson_residual_pair", None, i64, i64, i64, vp, vp, vp, i32p, i64p, i64, vp,
         C.POINTER(C.c_double), C.POINTER(C.c_double))
    _LIB = Oracle(lib)
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data


class Oracle:
    """Thin numpy-friendly wrappers over liborc.so."""

    def __init__(self, lib):
        self.lib = lib

    def __getattr__(self, name):
        return getattr(self.lib, name)

    def halo_fill(self, g, bid, halo, full, bc=1):
        hsx = g.B + 2 * halo
        tile = np.zeros(hsx ** 3, np.float32)
        self.lib.orc_halo_fill(g.sx, g.sy, g.sz, g.L, np.ascontiguousarray(g.slab), g.b2s, bid, halo, int(full),
                               bc, tile)
        return tile.reshape(hsx, hsx, hsx)

    def kernel(self, stencil, tile, dt, clamp01=False, B=16):
        out = np.zeros(B ** 3, np.float32)
        name = {0: "laplacian", 1: "meancurv", 2: "bilaplacian"}[stencil]
        getattr(self.lib, "orc_kernel_" + name)(B, np.ascontiguousarray(tile, np.float32).ravel(), dt,
                                                 int(clamp01), out)
        return out

    def sweep(self, g, active, stencil, dt, iterations, clamp01=False, nthreads=1, ftz=False):
        """In place over g's slab (Sweeper::sweep)."""
        slab = np.ascontiguousarray(g.slab)
        act = np.ascontiguousarray(active, np.int64)
        self.lib.orc_sweep(g.sx, g.sy, g.sz, slab, g.b2s, act, act.size, stencil, dt, iterations, int(clamp01),
                           nthreads, int(ftz))
        g._slab[: g.n_slots] = slab

    def color_buckets(self, g, active):
        act = np.ascontiguousarray(active, np.int64)
        out = np.zeros(max(act.size, 1), np.int64)
        counts = np.zeros(8, np.int64)
        self.lib.orc_color_buckets(g.sx, g.sy, g.sz, g.L, act, act.size, out, counts)
        return out[: act.size], counts

    def sample(self, g, pos, interp, align, bc, want_grad=False, want_hess=False):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        val = np.zeros(n, np.float32)
        grad = np.zeros((n, 3), np.float32) if want_grad else None
        hess = np.zeros((n, 6), np.float32) if want_hess else None
        self.lib.orc_sample(g.sx, g.sy, g.sz, np.ascontiguousarray(g.slab), g.b2s, pos, n, interp, align, bc,
                            _ptr(val), _ptr(grad), _ptr(hess))
        return val, grad, hess

    def active_blocks(self, sx, sy, sz, pos, r_scan):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        nb = ((sx + 15) >> 4) * ((sy + 15) >> 4) * ((sz + 15) >> 4)
        out = np.zeros(nb, np.int64)
        k = self.lib.orc_active_blocks(sx, sy, sz, pos, pos.shape[0], r_scan, out)
        return out[:k]

    def bucket_indices(self, bids, n_blocks):
        bids = np.ascontiguousarray(bids, np.int64)
        n = bids.size
        pb = np.zeros(max(n, 1), np.int64)
        starts = np.zeros(n + 1, np.int64)
        perm = np.zeros(max(n, 1), np.int64)
        k = self.lib.orc_bucket_indices(bids, n, n_blocks, pb, starts, perm)
        return pb[:k], starts[: k + 1], perm[:n]
