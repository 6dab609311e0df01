"""ctypes binding of the surrogate kernels (include/meep_surrogate.h) and a
thin strided-GEMM helper that picks the reduction split and owns the split
workspace."""
import ctypes
from ctypes import c_float, c_int, c_void_p

import torch

from .. import _engine as E

c_ll = ctypes.c_longlong


class Gemm(ctypes.Structure):
    _fields_ = [("M", c_int), ("N", c_int), ("K", c_int), ("R", c_int), ("batch", c_int), ("splits", c_int),
                ("A", c_void_p), ("a_m", c_ll), ("a_k", c_ll), ("a_b", c_ll), ("a_r", c_ll),
                ("B", c_void_p), ("b_k", c_ll), ("b_n", c_ll), ("b_b", c_ll), ("b_r", c_ll),
                ("C", c_void_p), ("c_m", c_ll), ("c_n", c_ll), ("c_b", c_ll),
                ("work", c_void_p), ("alpha", c_float), ("beta", c_float)]


def lib():
    L = E.lib()
    if getattr(L, "_nn_bound", False):
        return L
    P = ctypes.POINTER
    L.mt_gemm_bf16.argtypes = [P(Gemm), c_void_p]
    L.mt_cast_bf16.argtypes = [c_void_p, c_void_p, c_ll, c_void_p]
    L.mt_bias_act.argtypes = [c_void_p, c_void_p, c_int, c_int, c_ll, c_int, c_void_p, c_void_p, c_void_p]
    L.mt_act_grad.argtypes = [c_void_p, c_void_p, c_ll, c_int, c_int, c_int, c_float, c_void_p, c_void_p,
                              c_void_p]
    L.mt_channel_sum.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_float, c_void_p]
    L.mt_row_sum.argtypes = [c_void_p, c_void_p, c_ll, c_int, c_float, c_void_p]
    L.mt_pack_spectral.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]
    L.mt_unpack_spectral_grad.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]
    L.mt_adam.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_ll, c_float, c_float, c_float, c_float,
                          c_int, c_float, c_void_p]
    for n in ("mt_gemm_bf16", "mt_cast_bf16", "mt_bias_act", "mt_act_grad", "mt_channel_sum", "mt_row_sum",
              "mt_pack_spectral", "mt_unpack_spectral_grad", "mt_adam"):
        getattr(L, n).restype = c_int
    L._nn_bound = True
    return L


def _p(t):
    return c_void_p(t.data_ptr()) if t is not None else c_void_p(0)


TARGET_WG = 512     # workgroups one GEMM launch aims for (2 per CU)
PROFILE = None      # list -> every GEMM appends (flops, algorithmic bytes, start event, end event)
MIN_TILES = 8       # fewest 32-deep reduction tiles per split


class Workspace:
    """Grows on demand; one per SpectralOperator (same stream)."""

    def __init__(self, device):
        self.device = device
        self.buf = torch.empty(0, dtype=torch.float32, device=device)

    def get(self, n):
        if self.buf.numel() < n:
            self.buf = torch.empty(int(n * 1.25) + 1024, dtype=torch.float32, device=self.device)
        return self.buf


def gemm(A, a, B, b, C, c, M, N, K, *, batch=1, R=1, alpha=1.0, beta=0.0, ws=None, splits=None):
    """C[z] = alpha * sum_r A[z,r] B[z,r] + beta * C[z] on the matrix cores.

    a = (a_m, a_k, a_b, a_r), b = (b_k, b_n, b_b, b_r), c = (c_m, c_n, c_b):
    element strides of the BF16 operands A, B and the FP32 result C."""
    E.require_device(A, B, C)
    if A.dtype != torch.bfloat16 or B.dtype != torch.bfloat16 or C.dtype != torch.float32:
        raise TypeError("gemm: A, B must be bfloat16 and C float32")
    tiles = -(-M // 64) * -(-N // 64) * batch
    f_total = R * -(-K // 32)
    if splits is None:
        splits = 1
        if tiles < TARGET_WG and f_total >= 2 * MIN_TILES:
            splits = min(TARGET_WG // tiles, f_total // MIN_TILES, 65535 // batch)
            splits = max(splits, 1)
    g = Gemm(M, N, K, R, batch, splits, _p(A), *[int(x) for x in a], _p(B), *[int(x) for x in b],
             _p(C), *[int(x) for x in c], c_void_p(0), float(alpha), float(beta))
    if splits > 1:
        if ws is None:
            raise ValueError("gemm: a split reduction needs a workspace")
        g.work = _p(ws.get(splits * batch * M * N))
    if PROFILE is not None:
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        ev[0].record()
    E.check(lib().mt_gemm_bf16(ctypes.byref(g), E.stream_handle()), "mt_gemm_bf16")
    if PROFILE is not None:
        ev[1].record()
        nbytes = (2.0 * (M * K + K * N) * R + 4.0 * M * N * (2 if beta else 1)) * batch
        PROFILE.append((2.0 * M * N * K * R * batch, nbytes, ev[0], ev[1]))
    return C


def cast_bf16(x, out=None):
    E.require_device(x)
    if x.dtype != torch.float32 or not x.is_contiguous():
        raise TypeError("cast_bf16: needs a contiguous float32 tensor")
    if out is None:
        out = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
    E.check(lib().mt_cast_bf16(_p(x), _p(out), x.numel(), E.stream_handle()), "mt_cast_bf16")
    return out


def bias_act(x, bias, C, P, act, act16=None, act32=None):
    E.check(lib().mt_bias_act(_p(x), _p(bias), C, P, x.numel(), int(act), _p(act16), _p(act32),
                              E.stream_handle()), "mt_bias_act")


def act_grad(dy, pre, n, act, P=1, row_bcast=False, scale=1.0, d32=None, d16=None):
    E.check(lib().mt_act_grad(_p(dy), _p(pre), n, int(act), P, int(row_bcast), float(scale), _p(d32), _p(d16),
                              E.stream_handle()), "mt_act_grad")


def channel_sum(x, out, B, C, P, scale=1.0, ws=None):
    """out[c] = scale * sum over b, p of x[b, c, p]; the B*C row sums go to
    the workspace ws (a fresh buffer when None)."""
    work = ws.get(B * C) if ws is not None else torch.empty(B * C, dtype=torch.float32, device=x.device)
    E.check(lib().mt_channel_sum(_p(x), _p(out), _p(work), B, C, P, float(scale), E.stream_handle()),
            "mt_channel_sum")


def row_sum(x, out, rows, P, scale=1.0):
    E.check(lib().mt_row_sum(_p(x), _p(out), rows, P, float(scale), E.stream_handle()), "mt_row_sum")


def pack_spectral(wre, wim, wp, Ci, Co, Mo):
    E.check(lib().mt_pack_spectral(_p(wre), _p(wim), _p(wp), Ci, Co, Mo, E.stream_handle()), "mt_pack_spectral")


def unpack_spectral_grad(dwp, dwre, dwim, Ci, Co, Mo):
    E.check(lib().mt_unpack_spectral_grad(_p(dwp), _p(dwre), _p(dwim), Ci, Co, Mo, E.stream_handle()),
            "mt_unpack_spectral_grad")


def adam(w, g, m, v, lr, b1, b2, eps, step, wd=0.0):
    E.check(lib().mt_adam(_p(w), _p(g), _p(m), _p(v), w.numel(), float(lr), float(b1), float(b2), float(eps),
                          int(step), float(wd), E.stream_handle()), "mt_adam")
