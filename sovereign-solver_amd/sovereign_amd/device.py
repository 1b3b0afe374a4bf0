"""ctypes helpers over the sov_hip C-ABI (include/sov_hip.h): a context,
device buffers, CSR handles and event timing, for tests and benchmarks.  All
work runs in libsov_hip.so; nothing here computes."""
from __future__ import annotations

import ctypes as C

import numpy as np

import sovereign_amd as S  # noqa: E402  (package root: the library loaders)

_v = C.c_void_p
_i64 = C.c_int64


def lib():
    L = S.hip()
    if getattr(L, "_typed", False):
        return L
    d = C.c_double
    L.sov_hip_device_count.argtypes = [C.POINTER(C.c_int)]
    L.sov_hip_ctx_create.argtypes = [C.c_int, C.POINTER(_v)]
    L.sov_hip_ctx_destroy.argtypes = [_v]
    L.sov_hip_ctx_device_name.argtypes = [_v]
    L.sov_hip_malloc.argtypes = [_v, C.c_size_t, C.POINTER(_v)]
    L.sov_hip_free.argtypes = [_v, _v]
    for f in ("sov_hip_memcpy_h2d", "sov_hip_memcpy_d2h", "sov_hip_memcpy_d2d"):
        getattr(L, f).argtypes = [_v, _v, _v, C.c_size_t]
    L.sov_hip_fill_zero.argtypes = [_v, _v, C.c_size_t]
    L.sov_hip_sync.argtypes = [_v]
    L.sov_hip_axpy_f64.argtypes = [_v, _i64, d, _v, _v]
    L.sov_hip_scal_f64.argtypes = [_v, _i64, d, _v]
    L.sov_hip_dot_f64.argtypes = [_v, _i64, _v, _v, C.POINTER(d)]
    L.sov_hip_nrm2_f64.argtypes = [_v, _i64, _v, C.POINTER(d)]
    L.sov_hip_amax_f64.argtypes = [_v, _i64, _v, C.POINTER(d)]
    L.sov_hip_project_box_f64.argtypes = [_v, _i64, _v, _v, _v]
    L.sov_hip_convert_f64_f32.argtypes = [_v, _i64, _v, _v]
    L.sov_hip_convert_f32_f64.argtypes = [_v, _i64, _v, _v]
    L.sov_hip_csr_create.argtypes = [_v, C.c_int32, C.c_int32, C.c_int32, _v, _v, _v, C.c_int, C.POINTER(_v)]
    L.sov_hip_csr_destroy.argtypes = [_v]
    L.sov_hip_csr_make_f32.argtypes = [_v, _v]
    L.sov_hip_csr_scale.argtypes = [_v, _v, _v, _v]
    L.sov_hip_csr_transpose_arrays.argtypes = [_v, _v, _v, _v, _v]
    L.sov_hip_csr_set_lanes.argtypes = [_v, C.c_int, C.POINTER(C.c_int32)]
    L.sov_hip_csr_get_lanes.argtypes = [_v, C.c_int, C.POINTER(C.c_int32)]
    L.sov_hip_csr_tune.argtypes = [_v, _v, C.c_int]
    L.sov_hip_csr_set_segments.argtypes = [_v, _v, C.c_int, C.c_int]
    L.sov_hip_csr_get_segments.argtypes = [_v, C.c_int, C.POINTER(C.c_int)]
    L.sov_hip_csr_set_tiles.argtypes = [_v, _v, C.c_int, C.c_int]
    L.sov_hip_csr_get_tiles.argtypes = [_v, C.c_int, C.POINTER(C.c_int)]
    L.sov_hip_csr_tiles_usable.argtypes = [_v, _v, C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.sov_hip_csr_get_tiles_prec.argtypes = [_v, C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.sov_hip_tile_layout.argtypes = [C.c_int32, C.c_int32, C.c_int32, _v, _v, _v, C.c_int32, C.c_int32, _v, _v, _v, _v,
                                      _v, _v]
    L.sov_hip_tile_passes.argtypes = [C.c_int32, C.c_int32, C.c_int32, _v, _v, _v, C.c_int32, C.c_int32, _v, _v, _v]
    L.sov_hip_spmv.argtypes = [_v, _v, C.c_int, C.c_int, d, _v, d, _v]
    L.sov_hip_spmv_device.argtypes = [_v, _v, C.c_int, C.c_int, d, _v, d, _v]
    L.sov_hip_csr_order.argtypes = [_v, _v, C.c_int, _v]
    L.sov_hip_csr_permute.argtypes = [_v, _v, C.c_int, C.c_int, C.c_int, _v, _v]
    L.sov_hip_scale_ruiz_sweep.argtypes = [_v, _v, _v, _v, C.POINTER(d)]
    L.sov_hip_scale_pock_chambolle.argtypes = [_v, _v, d, _v, _v]
    L.sov_hip_timer_start.argtypes = [_v]
    L.sov_hip_timer_stop.argtypes = [_v, C.POINTER(d)]
    L.sov_hip_events_create.argtypes = [_v, C.c_int, C.POINTER(_v)]
    L.sov_hip_events_record.argtypes = [_v, _v, C.c_int]
    L.sov_hip_events_elapsed.argtypes = [_v, C.c_int, C.c_int, C.POINTER(d)]
    L.sov_hip_events_destroy.argtypes = [_v]
    L.sov_hip_pdhg_primal_step.argtypes = [_v, _v, _v, d, C.c_int]
    L.sov_hip_pdhg_dual_step.argtypes = [_v, _v, _v, d, d, C.c_int]
    L.sov_hip_pdhg_check.argtypes = [_v, _v, _v, _v]
    L.sov_hip_pdhg_check_static.argtypes = [_v, _v, _v, _v]
    L.sov_hip_pdhg_restart.argtypes = [_v, _v, C.c_int64, _v, _v, _v, C.POINTER(d)]
    L.sov_hip_pdhg_set_point.argtypes = [_v, _v, _v, _v, _v]
    L.sov_hip_pdhg_anchor.argtypes = [_v, _v, _v, _v, _v, _v, _v, d]
    L.sov_hip_pdhg_combine.argtypes = [_v, _v, _v, _v, _v, C.c_int64, d]
    L.sov_hip_graph_begin.argtypes = [_v]
    L.sov_hip_graph_end.argtypes = [_v, C.POINTER(_v)]
    L.sov_hip_graph_launch.argtypes = [_v, _v]
    L.sov_hip_graph_destroy.argtypes = [_v]
    L.sov_hip_ctx_props.argtypes = [_v, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    f32p = C.POINTER(C.c_float)
    L.sov_hip_mlp_create.argtypes = [_v, C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.POINTER(_v)]
    L.sov_hip_mlp_destroy.argtypes = [_v]
    L.sov_hip_mlp_num_params.argtypes = [_v]
    L.sov_hip_mlp_num_params.restype = C.c_int64
    L.sov_hip_mlp_set_params.argtypes = [_v, _v, _v]
    L.sov_hip_mlp_get_params.argtypes = [_v, _v, _v, _v]
    L.sov_hip_mlp_forward.argtypes = [_v, _v, C.c_int32, _v, _v]
    L.sov_hip_mlp_backward.argtypes = [_v, _v, _v]
    L.sov_hip_mlp_adam.argtypes = [_v, _v, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32]
    L.sov_hip_mlp_mse.argtypes = [_v, _v, C.c_int32, _v, _v, _v, _v, C.POINTER(d)]
    L._typed = True
    return L


def check(rc):
    if rc != 0:
        raise RuntimeError(f"sov_hip call failed ({rc}): {lib().sov_hip_last_error().decode()}")


class Ctx:
    def __init__(self, device=0):
        self.L = lib()
        self.h = _v()
        check(self.L.sov_hip_ctx_create(device, C.byref(self.h)))
        self._bufs = []

    def close(self):
        for p in self._bufs:
            self.L.sov_hip_free(self.h, p)
        self._bufs = []
        self.L.sov_hip_ctx_destroy(self.h)

    def alloc(self, nbytes):
        p = _v()
        check(self.L.sov_hip_malloc(self.h, max(nbytes, 8), C.byref(p)))
        self._bufs.append(p)
        return p

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        if a.nbytes:
            check(self.L.sov_hip_memcpy_h2d(self.h, p, a.ctypes.data, a.nbytes))
        return p

    def download(self, p, count, dtype=np.float64):
        out = np.zeros(count, dtype)
        if out.nbytes:
            check(self.L.sov_hip_memcpy_d2h(self.h, out.ctypes.data, p, out.nbytes))
        return out

    def free(self, p):
        self._bufs.remove(p)
        check(self.L.sov_hip_free(self.h, p))


class Csr:
    """A sov_hip_csr built from host CSR arrays (device copies owned here)."""

    def __init__(self, ctx: Ctx, m, n, row_ptr, col_idx, val, own_values=True):
        self.ctx, self.m, self.n = ctx, m, n
        self.nnz = int(row_ptr[-1])
        self.rp = ctx.upload(np.asarray(row_ptr, np.int32))
        self.ci = ctx.upload(np.asarray(col_idx, np.int32)[: max(self.nnz, 0)])
        self.v = ctx.upload(np.asarray(val, np.float64)[: max(self.nnz, 0)])
        self.h = _v()
        check(ctx.L.sov_hip_csr_create(ctx.h, m, n, self.nnz, self.rp, self.ci, self.v, 1 if own_values else 0,
                                       C.byref(self.h)))

    def destroy(self):
        self.ctx.L.sov_hip_csr_destroy(self.h)

    def order(self, which):
        """Device order of the rows (which = 0) or columns (1): perm[p] is the
        natural index stored at device position p."""
        k = self.m if which == 0 else self.n
        perm = np.zeros(max(k, 1), np.int32)
        check(self.ctx.L.sov_hip_csr_order(self.ctx.h, self.h, which, perm.ctypes.data))
        return perm[:k]

    def spmv(self, x, transpose=False, alpha=1.0, beta=0.0, y=None, prec=0):
        dt = np.float64 if prec == 0 else np.float32
        nout = self.n if transpose else self.m
        dx = self.ctx.upload(np.asarray(x, dt))
        dy = self.ctx.upload(np.zeros(nout, dt) if y is None else np.asarray(y, dt))
        check(self.ctx.L.sov_hip_spmv(self.ctx.h, self.h, int(transpose), prec, alpha, dx, beta, dy))
        out = self.ctx.download(dy, nout, dt)
        self.ctx.free(dx)
        self.ctx.free(dy)
        return out


class PdhgState(C.Structure):
    """sov_hip_pdhg_state (include/sov_hip.h)."""
    _fields_ = [("prec", C.c_int32), ("m", C.c_int32), ("n", C.c_int32), ("c", _v), ("col_lo", _v),
                ("col_hi", _v), ("row_lo", _v), ("row_hi", _v), ("x", _v), ("x_bar", _v), ("x_sum", _v), ("s", _v),
                ("y", _v), ("y_sum", _v), ("s_grad", _v)]


class CheckArgs(C.Structure):
    _fields_ = [("st", C.POINTER(PdhgState)), ("since_restart", C.c_int64), ("tol_scaled", C.c_double),
                ("tol_orig", C.c_double), ("obj_sign", C.c_double), ("obj_const", C.c_double), ("row_scale", _v),
                ("col_scale", _v), ("orig_row_lo", _v), ("orig_row_hi", _v), ("orig_col_lo", _v),
                ("orig_col_hi", _v), ("orig_c", _v), ("act_avg", _v), ("base", C.POINTER(PdhgState)),
                ("base_x", _v), ("base_y", _v), ("base_act", _v), ("base_aty", _v), ("inv_delta", C.c_double)]


class CheckOut(C.Structure):
    _fields_ = [("kkt_cur", C.c_double), ("kkt_avg", C.c_double), ("orig_cur", C.c_double),
                ("orig_avg", C.c_double), ("rel_cur", C.c_double * 3), ("rel_avg", C.c_double * 3),
                ("pobj_cur", C.c_double), ("dobj_cur", C.c_double), ("pobj_avg", C.c_double),
                ("dobj_avg", C.c_double)]


class Events:
    def __init__(self, ctx: Ctx, count: int):
        self.ctx = ctx
        self.h = _v()
        check(ctx.L.sov_hip_events_create(ctx.h, count, C.byref(self.h)))

    def record(self, i):
        check(self.ctx.L.sov_hip_events_record(self.ctx.h, self.h, i))

    def elapsed(self, i, j):
        d = C.c_double()
        check(self.ctx.L.sov_hip_events_elapsed(self.h, i, j, C.byref(d)))
        return d.value

    def destroy(self):
        self.ctx.L.sov_hip_events_destroy(self.h)
