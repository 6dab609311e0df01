"""ctypes binding to libmeepfdtd.so (the C ABI declared in include/meep_fdtd.h).

Torch owns every device buffer; this module only passes raw device pointers,
sizes and the current HIP stream.  There is no CPU path: if the library is
missing or a tensor is not on a HIP device, the call raises.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmeepfdtd.so")

MT_F32, MT_F64 = 0, 1
MAX_TERMS = 4
ABI_VERSION = 27

c_int, c_ll, c_double, c_void_p = ctypes.c_int, ctypes.c_longlong, ctypes.c_double, ctypes.c_void_p


class PmlAxis(ctypes.Structure):
    _fields_ = [("b", c_void_p), ("a", c_void_p), ("n", c_int), ("lo_end", c_int),
                ("hi_begin", c_int)]


class Source2D(ctypes.Structure):
    _fields_ = [("present", c_int), ("i0", c_int), ("i1", c_int), ("j0", c_int), ("j1", c_int),
                ("w", c_void_p), ("wstep", c_ll)]


class TM2D(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("dtype", c_int), ("y_pmc", c_int),
                ("dt", c_double), ("dx", c_double), ("eps", c_void_p),
                ("hy", PmlAxis), ("hx", PmlAxis), ("ex", PmlAxis), ("ey", PmlAxis),
                ("Ez", c_void_p), ("Dz", c_void_p), ("Hx", c_void_p), ("Hy", c_void_p),
                ("psi_Ezy", c_void_p), ("psi_Ezx", c_void_p),
                ("psi_Hyx", c_void_p), ("psi_Hxy", c_void_p), ("tape", c_void_p),
                ("variant", c_int), ("pad_", c_int), ("alt", c_void_p * 8),
                ("hsrc", Source2D * 2), ("hcoef", ctypes.POINTER(c_double)),
                ("n_poles", c_int), ("kerr", c_int), ("c1", c_void_p), ("c2", c_void_p), ("c3", c_void_p),
                ("P_cur", c_void_p), ("P_prev", c_void_p), ("chi2", c_void_p), ("chi3", c_void_p),
                ("d1", c_void_p), ("d2", c_void_p), ("h1x", c_void_p), ("h2x", c_void_p),
                ("h1y", c_void_p), ("h2y", c_void_p),
                ("n_mpoles", c_int), ("pad4_", c_int), ("mc1x", c_void_p), ("mc2x", c_void_p),
                ("mc3x", c_void_p), ("mc1y", c_void_p), ("mc2y", c_void_p), ("mc3y", c_void_p),
                ("Bx", c_void_p), ("By", c_void_p), ("Mx_cur", c_void_p), ("Mx_prev", c_void_p),
                ("My_cur", c_void_p), ("My_prev", c_void_p),
                ("y_mirror", c_int), ("x_mirror", c_int), ("y_rot2", c_int), ("rot4", c_int),
                ("Hx0", c_void_p), ("Hy0", c_void_p)]


class TM2DAdj(ctypes.Structure):
    _fields_ = [("aEz", c_void_p), ("aDz", c_void_p), ("aHx", c_void_p), ("aHy", c_void_p),
                ("apsi_Ezy", c_void_p), ("apsi_Ezx", c_void_p), ("apsi_Hyx", c_void_p),
                ("apsi_Hxy", c_void_p), ("gHy", c_void_p), ("gHx", c_void_p), ("gEx", c_void_p),
                ("gEy", c_void_p), ("eps_grad", c_void_p), ("aP_cur", c_void_p), ("aP_prev", c_void_p),
                ("chi2_grad", c_void_p), ("chi3_grad", c_void_p), ("c3_grad", c_void_p), ("ez_init", c_void_p),
                ("aBx", c_void_p), ("aBy", c_void_p), ("aMx_cur", c_void_p), ("aMx_prev", c_void_p),
                ("aMy_cur", c_void_p), ("aMy_prev", c_void_p)]


class TE2D(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("dtype", c_int), ("pad_", c_int),
                ("dt", c_double), ("dx", c_double), ("eps_x", c_void_p), ("eps_y", c_void_p),
                ("hx", PmlAxis), ("hy", PmlAxis), ("ey", PmlAxis), ("ex", PmlAxis),
                ("Hz", c_void_p), ("Ex", c_void_p), ("Ey", c_void_p), ("Dx", c_void_p),
                ("Dy", c_void_p), ("psi_Eyx", c_void_p), ("psi_Exy", c_void_p),
                ("psi_Hzy", c_void_p), ("psi_Hzx", c_void_p),
                ("variant", c_int), ("pad2_", c_int), ("alt", c_void_p * 8), ("tape", c_void_p),
                ("hsrc", Source2D), ("hcoef", ctypes.POINTER(c_double)),
                ("npx", c_int), ("npy", c_int), ("c1x", c_void_p), ("c2x", c_void_p), ("c3x", c_void_p),
                ("c1y", c_void_p), ("c2y", c_void_p), ("c3y", c_void_p), ("Px_cur", c_void_p),
                ("Px_prev", c_void_p), ("Py_cur", c_void_p), ("Py_prev", c_void_p), ("d1x", c_void_p),
                ("d2x", c_void_p), ("d1y", c_void_p), ("d2y", c_void_p), ("hz1", c_void_p), ("hz2", c_void_p),
                ("kerr", c_int), ("pad3_", c_int), ("chi2x", c_void_p), ("chi3x", c_void_p), ("chi2y", c_void_p),
                ("chi3y", c_void_p), ("kx", c_void_p), ("ky", c_void_p),
                ("inv_xx_x", c_void_p), ("inv_xy_x", c_void_p), ("inv_yy_y", c_void_p), ("inv_xy_y", c_void_p),
                ("c3xy", c_void_p), ("c3yx", c_void_p), ("y_mirror", c_int), ("x_mirror", c_int)]


class TE2DAdj(ctypes.Structure):
    _fields_ = [("aHz", c_void_p), ("aEx", c_void_p), ("aEy", c_void_p), ("aDx", c_void_p),
                ("aDy", c_void_p), ("apsi_Eyx", c_void_p), ("apsi_Exy", c_void_p),
                ("apsi_Hzy", c_void_p), ("apsi_Hzx", c_void_p), ("gEy", c_void_p), ("gEx", c_void_p),
                ("gDx", c_void_p), ("gDy", c_void_p), ("epsx_grad", c_void_p), ("epsy_grad", c_void_p),
                ("aPx_cur", c_void_p), ("aPx_prev", c_void_p), ("aPy_cur", c_void_p), ("aPy_prev", c_void_p),
                ("kerr_ws", c_void_p), ("ixx_grad", c_void_p), ("ixy_x_grad", c_void_p),
                ("iyy_grad", c_void_p), ("ixy_y_grad", c_void_p)]


class Yee3D(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("dtype", c_int), ("lean", c_int),
                ("absorber", c_int), ("dt", c_double), ("dx", c_double), ("eps", c_void_p),
                ("h", PmlAxis * 3), ("e", PmlAxis * 3),
                ("Ex", c_void_p), ("Ey", c_void_p), ("Ez", c_void_p),
                ("Hx", c_void_p), ("Hy", c_void_p), ("Hz", c_void_p),
                ("psi", c_void_p * 12), ("tape", c_void_p), ("variant", c_int), ("pad2_", c_int),
                ("alt", c_void_p * 18), ("damp", c_void_p * 12), ("n_poles", c_int), ("pad3_", c_int),
                ("c1", c_void_p), ("c2", c_void_p), ("c3", c_void_p * 3), ("P_cur", c_void_p * 3),
                ("P_prev", c_void_p * 3)]


class Yee3DAdj(ctypes.Structure):
    _fields_ = [("aEx", c_void_p), ("aEy", c_void_p), ("aEz", c_void_p), ("aHx", c_void_p),
                ("aHy", c_void_p), ("aHz", c_void_p), ("apsi", c_void_p * 12), ("g", c_void_p * 12),
                ("geps", c_void_p * 3), ("eps_grad", c_void_p), ("aP_cur", c_void_p * 3),
                ("aP_prev", c_void_p * 3)]


class Source3D(ctypes.Structure):
    _fields_ = [("kind", c_int), ("comp", c_int), ("i0", c_int), ("i1", c_int), ("j0", c_int),
                ("j1", c_int), ("k0", c_int), ("k1", c_int), ("w", c_void_p)]


class DftTerm(ctypes.Structure):
    _fields_ = [("a", c_void_p), ("b", c_void_p), ("s0", c_ll), ("s1", c_ll), ("t_half", c_int),
                ("field", c_int), ("line", c_int), ("fix", c_int), ("lo", c_int), ("pad_", c_int)]


class Monitor(ctypes.Structure):
    _fields_ = [("kind", c_int), ("n_terms", c_int), ("n0", c_int), ("n1", c_int),
                ("nfreq", c_int), ("pad_", c_int), ("term", DftTerm * MAX_TERMS),
                ("omega", c_void_p), ("acc", c_void_p)]


_lib = None


def lib():
    """Load libmeepfdtd.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"meep_torch_amd: HIP engine library not found at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    L.mt_abi_version.restype = c_int
    L.mt_last_error.restype = ctypes.c_char_p
    L.mt_stage_bytes.restype = c_ll
    L.mt_stage_bytes.argtypes = [P(Monitor), c_int, c_int, c_int]
    common = [P(Monitor), c_int, c_void_p, c_ll, c_int, c_int, c_int, c_void_p]
    L.mt_tm2d_run.restype = c_int
    L.mt_tm2d_run.argtypes = [P(TM2D), P(Source2D), P(c_double)] + common
    L.mt_te2d_run.restype = c_int
    L.mt_te2d_run.argtypes = [P(TE2D), P(Source2D), P(c_double)] + common
    L.mt_yee3d_run.restype = c_int
    L.mt_yee3d_run.argtypes = [P(Yee3D), P(Source3D), P(c_double)] + common
    L.mt_tiled_variants.restype = c_int
    L.mt_tiled_variants.argtypes = [c_int, P(c_int), P(c_int), P(c_int)]
    L.mt_tm2d_adjoint.restype = c_int
    L.mt_tm2d_adjoint.argtypes = [P(TM2D), P(TM2DAdj), c_void_p, P(Monitor), c_int, c_int, c_int,
                                  c_void_p]
    L.mt_te2d_adjoint.restype = c_int
    L.mt_te2d_adjoint.argtypes = [P(TE2D), P(TE2DAdj), c_void_p, P(Monitor), c_int, c_int, c_int,
                                  c_void_p]
    L.mt_kernel_timing_enable.restype = c_int
    L.mt_kernel_timing_enable.argtypes = [c_int]
    L.mt_kernel_timing_read.restype = c_int
    L.mt_kernel_timing_read.argtypes = [P(c_double), P(c_ll), P(c_ll)]
    L.mt_yee3d_adjoint.restype = c_int
    L.mt_yee3d_adjoint.argtypes = [P(Yee3D), P(Yee3DAdj), c_void_p, P(Monitor), c_int, c_int, c_int,
                                   c_void_p]
    L.mt_abi_layout.restype = c_int
    L.mt_abi_layout.argtypes = [P(c_ll), c_int]
    L.mt_tm2d_tiles.restype = c_int
    L.mt_tm2d_tiles.argtypes = [P(TM2D), P(c_int), P(c_int)]
    L.mt_te2d_tiles.restype = c_int
    L.mt_te2d_tiles.argtypes = [P(TE2D), P(c_int), P(c_int)]
    L.mt_tm2d_run_static.restype = c_int
    L.mt_tm2d_run_static.argtypes = [P(TM2D), P(Source2D), P(c_double), P(Monitor), c_int, c_int, c_int, c_void_p]
    L.mt_static_stage_bytes.restype = c_ll
    L.mt_div_recip_f32.restype = ctypes.c_float
    L.mt_div_recip_f32.argtypes = [ctypes.c_float]
    if L.mt_abi_version() != ABI_VERSION:
        raise RuntimeError("meep_torch_amd: libmeepfdtd.so ABI mismatch; rebuild")
    _lib = L
    return L


def check(rc, what):
    if rc != 0:
        msg = lib().mt_last_error().decode()
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def dtype_code(dtype):
    if dtype == torch.float32:
        return MT_F32
    if dtype == torch.float64:
        return MT_F64
    raise NotImplementedError(f"engine dtype must be float32 or float64, got {dtype}")


def require_device(*tensors):
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise RuntimeError(
                "meep_torch_amd runs on the HIP engine only: tensors must live on a ROCm device "
                "(device='cuda'); got a tensor on " + str(t.device))


def ptr(t):
    return c_void_p(t.data_ptr()) if t is not None else c_void_p(0)


def stream_handle():
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def shell_bounds(b, a):
    """(lo_end, hi_begin) of the CPML shells of a 1D profile: indices outside
    [0, lo_end) and [hi_begin, n) hold exactly b = 1, a = 0."""
    bc, ac = b.detach().reshape(-1).cpu(), a.detach().reshape(-1).cpu()
    n = bc.numel()
    active = ((bc != 1) | (ac != 0)).nonzero().flatten().tolist()
    if not active:
        return 0, n
    half = n // 2
    lo = [i for i in active if i < half]
    hi = [i for i in active if i >= half]
    lo_end = (max(lo) + 1) if lo else 0
    hi_begin = min(hi) if hi else n
    if lo_end > hi_begin:
        return n, n
    return lo_end, hi_begin


def pml_axis(b, a, bounds=None):
    b = b.reshape(-1)
    a = a.reshape(-1)
    lo, hi = bounds if bounds is not None else shell_bounds(b, a)
    return PmlAxis(ptr(b), ptr(a), b.numel(), lo, hi)


def mask_box(mask):
    """Bounding box of the nonzeros of a source mask (any rank) and the dense
    weights inside it; None for an all-zero mask."""
    nz = mask.detach().ne(0)
    if not bool(nz.any()):
        return None
    idx = nz.nonzero()
    lo = idx.min(0).values.tolist()
    hi = (idx.max(0).values + 1).tolist()
    sl = tuple(slice(l, h) for l, h in zip(lo, hi))
    return lo, hi, mask.detach()[sl].contiguous()


def view2d(t):
    """(data_ptr, s0, s1, n0, n1) of a 1D/2D strided view."""
    if t.dim() == 1:
        t = t.unsqueeze(0)
    if t.dim() != 2:
        raise ValueError("monitor view must be 1D or 2D")
    return t.data_ptr(), t.stride(0), t.stride(1), t.shape[0], t.shape[1]
