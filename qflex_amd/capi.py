"""ctypes binding of libqfx.so (the C-ABI declared in include/qfx.h).

Used by the tests and the benchmark to call the GPU primitives directly.
Raises if the library is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libqfx.so")

FP32, SPLIT_BF16, FP32_EXACT, SPLIT_BF16X3, SPLIT_FP16, FP8 = 0, 1, 2, 3, 4, 5
PRECISIONS = {"fp32": FP32, "split_bf16": SPLIT_BF16, "fp32_exact": FP32_EXACT,
              "split_bf16x3": SPLIT_BF16X3, "split_fp16": SPLIT_FP16, "fp8": FP8}
K_PERMUTE, K_CGEMM, K_GEMV, K_DOT, K_ACCUMULATE, K_COPY, K_CGEMM_SMALL = range(7)
TILE_256x128, TILE_128x256, TILE_512x32 = 1, 2, 3
FORM_ROW_MAJOR, FORM_A_ALONG_K, FORM_A_ALONG_M = 0, 1, 2
KERNEL_CLASSES = ["permute", "cgemm", "gemv", "dot", "accumulate", "copy", "cgemm_small"]
MLP_IN, MLP_HIDDEN, MLP_PARAMS = 64, 256, 82689

_c = ctypes
_vp = _c.c_void_p
_u64 = _c.c_uint64


class IndexMap(ctypes.Structure):
    """qfx_index_map: digits of a matrix index, least significant first."""
    _fields_ = [("ndigits", ctypes.c_int), ("bits", ctypes.c_ubyte * 16),
                ("shift", ctypes.c_ubyte * 16)]


class ScaleIO(ctypes.Structure):
    """qfx_scale_io: split-FP16 partial maxima passed along a chain of GEMMs."""
    _fields_ = [("a_max", ctypes.c_void_p), ("a_max_n", ctypes.c_int),
                ("b_max", ctypes.c_void_p), ("b_max_n", ctypes.c_int),
                ("c_max", ctypes.c_void_p), ("c_max_cap", ctypes.c_int),
                ("c_max_n", ctypes.c_int)]


def index_map(dims, axes):
    """Map of the matrix index over tensor axes `axes` (most significant
    first) of a row-major tensor with power-of-two `dims`."""
    lg = [int(d).bit_length() - 1 for d in dims]
    if any(1 << b != int(d) for b, d in zip(lg, dims)):
        raise ValueError("index_map needs power-of-two dimensions")
    strides = [sum(lg[a + 1:]) for a in range(len(dims))]
    m = IndexMap()
    digits = [a for a in reversed(list(axes)) if lg[a] > 0]
    if len(digits) > 16:
        raise ValueError("at most 16 digits")
    m.ndigits = len(digits)
    for j, a in enumerate(digits):
        m.bits[j] = lg[a]
        m.shift[j] = strides[a]
    return m

# name -> (restype, argtypes)
SIGNATURES = {
    "qfx_last_error": (_c.c_char_p, []),
    "qfx_context_create": (_c.c_int, [_c.c_int, _c.c_size_t, _c.POINTER(_vp)]),
    "qfx_context_destroy": (_c.c_int, [_vp]),
    "qfx_alloc": (_c.c_int, [_vp, _c.c_size_t, _c.POINTER(_vp)]),
    "qfx_free": (_c.c_int, [_vp, _vp]),
    "qfx_arena_stats": (_c.c_int, [_vp, _c.POINTER(_c.c_size_t), _c.POINTER(_c.c_size_t)]),
    "qfx_copy_h2d": (_c.c_int, [_vp, _vp, _vp, _c.c_size_t]),
    "qfx_copy_d2h": (_c.c_int, [_vp, _vp, _vp, _c.c_size_t]),
    "qfx_copy_d2d": (_c.c_int, [_vp, _vp, _vp, _c.c_size_t]),
    "qfx_zero": (_c.c_int, [_vp, _vp, _c.c_size_t]),
    "qfx_synchronize": (_c.c_int, [_vp]),
    "qfx_reorder_c64": (_c.c_int, [_vp, _vp, _vp, _c.c_int, _c.POINTER(_u64),
                                   _c.POINTER(_c.c_int)]),
    "qfx_cgemm_c64": (_c.c_int, [_vp, _c.c_int, _u64, _u64, _u64, _vp, _vp, _vp]),
    "qfx_cgemm_noalloc_c64": (_c.c_int, [_vp, _c.c_int, _u64, _u64, _u64, _vp, _vp, _vp]),
    "qfx_cgemm_mapped_c64": (_c.c_int, [_vp, _c.c_int, _vp, _c.POINTER(IndexMap),
                                        _c.POINTER(IndexMap), _vp, _c.POINTER(IndexMap),
                                        _c.POINTER(IndexMap), _vp]),
    "qfx_cgemm_mapped_ex_c64": (_c.c_int, [_vp, _c.c_int, _vp, _c.POINTER(IndexMap),
                                           _c.POINTER(IndexMap), _vp, _c.POINTER(IndexMap),
                                           _c.POINTER(IndexMap), _vp, _c.POINTER(ScaleIO)]),
    "qfx_cgemv_c64": (_c.c_int, [_vp, _c.c_int, _u64, _u64, _vp, _vp, _vp]),
    "qfx_cdotu_c64": (_c.c_int, [_vp, _u64, _vp, _vp, _vp]),
    "qfx_accumulate_c128": (_c.c_int, [_vp, _vp, _u64, _vp, _u64]),
    "qfx_accumulate_batched_c128": (_c.c_int, [_vp, _vp, _u64, _u64, _vp, _u64]),
    "qfx_mlp_forward": (_c.c_int, [_vp, _u64, _vp, _vp, _vp]),
    "qfx_mlp_backward": (_c.c_int, [_vp, _u64, _vp, _vp, _vp, _vp, _vp]),
    "qfx_mlp_adam": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _c.c_int, _c.c_float, _c.c_float,
                                _c.c_float, _c.c_float]),
    "qfx_profile_enable": (_c.c_int, [_vp, _c.c_int]),
    "qfx_profile_reset": (_c.c_int, [_vp]),
    "qfx_profile_read": (_c.c_int, [_vp, _c.POINTER(_u64), _c.POINTER(_c.c_double),
                                    _c.POINTER(_c.c_double), _c.POINTER(_c.c_double)]),
    "qfx_cgemm_grouped_c64": (_c.c_int, [_vp, _c.c_int, _c.c_int, _vp]),
    "qfx_copy_h2d_async": (_c.c_int, [_vp, _vp, _vp, _c.c_size_t]),
    "qfx_project_c64": (_c.c_int, [_vp, _vp, _vp, _u64, _u64, _u64, _u64, _c.POINTER(_u64)]),
    "qfx_amax_partials_c64": (_c.c_int, [_vp, _vp, _u64, _vp, _c.c_int, _c.POINTER(_c.c_int)]),
    "qfx_graph_begin": (_c.c_int, [_vp, _c.c_size_t]),
    "qfx_graph_end": (_c.c_int, [_vp, _c.POINTER(_vp)]),
    "qfx_graph_launch": (_c.c_int, [_vp, _vp]),
    "qfx_graph_destroy": (_c.c_int, [_vp]),
    "qfx_autotune_enable": (_c.c_int, [_vp, _c.c_int]),
    "qfx_autotune_entries": (_c.c_int, [_vp, _c.POINTER(_u64), _c.c_int, _c.POINTER(_c.c_int)]),
    "qfx_autotune_set": (_c.c_int, [_vp, _c.c_int, _c.c_int, _u64, _u64, _u64, _c.c_int]),
    "qfx_timer_start": (_c.c_int, [_vp]),
    "qfx_timer_stop": (_c.c_int, [_vp, _c.POINTER(_c.c_float)]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class QfxError(RuntimeError):
    pass


def check(status):
    if status != 0:
        raise QfxError(f"qfx status {status}: {load().qfx_last_error().decode()}")


class GemmDesc(_c.Structure):
    """qfx_gemm_desc: one product of a grouped launch."""
    _fields_ = [("m", _u64), ("n", _u64), ("k", _u64), ("A", _vp), ("B", _vp), ("C", _vp)]


class Graph:
    """A recorded sequence of a context's calls (qfx_graph_*)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle

    def launch(self):
        check(load().qfx_graph_launch(self.ctx.handle, self.handle))

    def destroy(self):
        if self.handle:
            check(load().qfx_graph_destroy(self.handle))
            self.handle = None


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _vp()
        check(load().qfx_alloc(ctx.handle, max(self.nbytes, 1), ctypes.byref(p)))
        self.ptr = p.value

    def free(self):
        if self.ptr:
            check(load().qfx_free(self.ctx.handle, self.ptr))
            self.ptr = None


class Context:
    """One qfx context (device stream + arena)."""

    def __init__(self, device=0, chunk_bytes=0):
        h = _vp()
        check(load().qfx_context_create(device, chunk_bytes, ctypes.byref(h)))
        self.handle = h.value

    def close(self):
        if self.handle:
            check(load().qfx_context_destroy(self.handle))
            self.handle = None

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        buf = self.alloc(arr.nbytes)
        check(load().qfx_copy_h2d(self.handle, buf.ptr, arr.ctypes.data, arr.nbytes))
        return buf

    def upload_async(self, buf, arr):
        """Staged host-to-device copy of `arr` into `buf` (returns before the copy is done)."""
        arr = np.ascontiguousarray(arr)
        check(load().qfx_copy_h2d_async(self.handle, buf.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, buf, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        check(load().qfx_copy_d2h(self.handle, out.ctypes.data, buf.ptr, out.nbytes))
        return out

    def copy(self, dst, src, nbytes):
        check(load().qfx_copy_d2d(self.handle, dst.ptr, src.ptr, nbytes))

    def arena_stats(self):
        """(bytes reserved from HIP, bytes handed out) of the context's arena."""
        r, u = _c.c_size_t(), _c.c_size_t()
        check(load().qfx_arena_stats(self.handle, _c.byref(r), _c.byref(u)))
        return r.value, u.value

    def sync(self):
        check(load().qfx_synchronize(self.handle))

    # primitives
    def reorder(self, src, dst, dims, perm):
        r = len(dims)
        d = (ctypes.c_uint64 * max(r, 1))(*dims)
        p = (ctypes.c_int * max(r, 1))(*perm)
        check(load().qfx_reorder_c64(self.handle, src.ptr, dst.ptr, r, d, p))

    def cgemm(self, precision, m, n, k, A, B, C):
        check(load().qfx_cgemm_c64(self.handle, precision, m, n, k, A.ptr, B.ptr, C.ptr))

    def cgemm_noalloc(self, precision, m, n, k, A, B, C):
        """qfx_cgemm_noalloc_c64: the same product, scratch from the fixed array."""
        check(load().qfx_cgemm_noalloc_c64(self.handle, precision, m, n, k, A.ptr, B.ptr, C.ptr))

    def cgemm_mapped(self, precision, A, a_rows, a_cols, B, b_rows, b_cols, C):
        """Returns False (C untouched) when no kernel reads this form."""
        st = load().qfx_cgemm_mapped_c64(self.handle, precision, A.ptr, ctypes.byref(a_rows),
                                         ctypes.byref(a_cols), B.ptr, ctypes.byref(b_rows),
                                         ctypes.byref(b_cols), C.ptr)
        if st == 4:  # QFX_ERR_UNSUPPORTED
            return False
        check(st)
        return True

    # ---- the cost model's network (QFX_MLP_PARAMS FP32 parameters)
    def mlp_forward(self, batch, X, params, y):
        check(load().qfx_mlp_forward(self.handle, batch, X.ptr, params.ptr, y.ptr))

    def mlp_backward(self, batch, X, target, params, grads, loss):
        check(load().qfx_mlp_backward(self.handle, batch, X.ptr, target.ptr, params.ptr,
                                      grads.ptr, loss.ptr))

    def mlp_adam(self, params, grads, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        check(load().qfx_mlp_adam(self.handle, params.ptr, grads.ptr, m.ptr, v.ptr, step, lr,
                                  beta1, beta2, eps))

    def cgemm_grouped(self, precision, problems):
        """problems: [(m, n, k, A, B, C)] with DeviceBuffers; one grouped call."""
        arr = (GemmDesc * max(len(problems), 1))()
        for i, (m, n, k, A, B, C) in enumerate(problems):
            arr[i] = GemmDesc(m, n, k, A.ptr, B.ptr, C.ptr)
        check(load().qfx_cgemm_grouped_c64(self.handle, precision, len(problems),
                                           _c.cast(arr, _vp)))

    def project(self, src, dst, pre, dim, post, picks):
        """dst[a][j][r] = src[a][picks[j]][r] (complex64 tensors as [pre][dim][post])."""
        arr = (_u64 * max(len(picks), 1))(*picks)
        check(load().qfx_project_c64(self.handle, src.ptr, dst.ptr, pre, dim, post, len(picks),
                                     arr))

    def amax_partials(self, X, n, out, cap=256):
        """Partial maxima of |re|, |im| of X[0..n) into device floats `out`; returns their count."""
        count = _c.c_int(0)
        check(load().qfx_amax_partials_c64(self.handle, X.ptr, n, out.ptr, cap, _c.byref(count)))
        return count.value

    def zero(self, buf, nbytes=None):
        check(load().qfx_zero(self.handle, buf.ptr, buf.nbytes if nbytes is None else nbytes))

    # ---- recorded sequences (HIP graphs)
    def graph_begin(self, workspace_bytes=0):
        check(load().qfx_graph_begin(self.handle, workspace_bytes))

    def graph_end(self):
        """Returns a Graph; replay with graph.launch()."""
        g = _vp()
        check(load().qfx_graph_end(self.handle, _c.byref(g)))
        return Graph(self, g.value)

    # ---- tile autotuner
    def autotune(self, on):
        check(load().qfx_autotune_enable(self.handle, 1 if on else 0))

    def autotune_entries(self):
        """[(precision, form, m, n, k, tile)] held by the context."""
        count = _c.c_int(0)
        check(load().qfx_autotune_entries(self.handle, None, 0, _c.byref(count)))
        buf = (_u64 * (6 * max(count.value, 1)))()
        check(load().qfx_autotune_entries(self.handle, buf, count.value, _c.byref(count)))
        return [tuple(int(buf[6 * i + j]) for j in range(6)) for i in range(count.value)]

    def autotune_set(self, precision, form, m, n, k, tile):
        check(load().qfx_autotune_set(self.handle, precision, form, m, n, k, tile))

    def cgemm_mapped_ex(self, precision, A, a_rows, a_cols, B, b_rows, b_cols, C, io):
        st = load().qfx_cgemm_mapped_ex_c64(self.handle, precision, A.ptr, ctypes.byref(a_rows),
                                            ctypes.byref(a_cols), B.ptr, ctypes.byref(b_rows),
                                            ctypes.byref(b_cols), C.ptr, ctypes.byref(io))
        if st == 4:
            return False
        check(st)
        return True

    def cgemv(self, trans, m, k, A, x, y):
        check(load().qfx_cgemv_c64(self.handle, trans, m, k, A.ptr, x.ptr, y.ptr))

    def cdotu(self, n, x, y, out):
        check(load().qfx_cdotu_c64(self.handle, n, x.ptr, y.ptr, out.ptr))

    def accumulate(self, src, n, amps, offset=0):
        check(load().qfx_accumulate_c128(self.handle, src.ptr, n, amps.ptr, offset))

    def accumulate_batched(self, src, n, nbatch, amps, offset=0):
        check(load().qfx_accumulate_batched_c128(self.handle, src.ptr, n, nbatch, amps.ptr,
                                                 offset))

    def profile(self, on=True):
        check(load().qfx_profile_enable(self.handle, 1 if on else 0))

    def profile_reset(self):
        check(load().qfx_profile_reset(self.handle))

    def profile_read(self):
        n = len(KERNEL_CLASSES)
        la = (ctypes.c_uint64 * n)()
        ms = (ctypes.c_double * n)()
        by = (ctypes.c_double * n)()
        fl = (ctypes.c_double * n)()
        check(load().qfx_profile_read(self.handle, la, ms, by, fl))
        return {KERNEL_CLASSES[i]: {"launches": la[i], "ms": ms[i], "bytes": by[i],
                                    "flops": fl[i]} for i in range(n)}

    def timer_start(self):
        check(load().qfx_timer_start(self.handle))

    def timer_stop(self):
        t = ctypes.c_float()
        check(load().qfx_timer_stop(self.handle, ctypes.byref(t)))
        return t.value
