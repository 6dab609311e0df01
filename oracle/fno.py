"""CPU oracle of the spectral neural-operator surrogate -- TEST INFRASTRUCTURE.

The surrogate has no counterpart in the reference (its oracle for the
physics stays the FDTD path); this module is the CPU restatement that the
HIP training and inference kernels are checked against: a float64 numpy
forward and a hand-derived backward of the same network, with optional
emulation of the BF16 rounding the GPU applies to every matrix-product
operand (FP32 value rounded to nearest-even BF16).  Sums and products are
float64 here, so with emulation on the only differences left are the GPU's
FP32 accumulation order and the rare BF16 rounding flips it causes.

Pinning (tests/test_oracle_fno.py): the truncated-DFT bases against
numpy.fft.rfft2 / irfft2; forward and parameter/input gradients against an
independent torch formulation built on torch.fft.rfft2 / irfft2 and
torch.autograd (float64); the backward against central finite differences;
the GELU derivative against finite differences.
"""
import math

import numpy as np

SQRT1_2 = 1.0 / math.sqrt(2.0)


def bf16(x):
    """Round to FP32, then to BF16 (nearest, ties to even); returns float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _erf(x):
    from scipy.special import erf
    return erf(x)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x * SQRT1_2))


def gelu_d(x):
    return 0.5 * (1.0 + _erf(x * SQRT1_2)) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def modes(n, mx, my):
    ks = [a if a < mx // 2 else n - mx + a for a in range(mx)]
    return [(kx, ky) for kx in ks for ky in range(my)]


def bases(n, mx, my):
    """F, G (2Mo, n*n): [Re, Im] rows of the ortho rfft2 restricted to the
    retained modes, and the matching irfft2 synthesis rows (ky > 0 doubled)."""
    md = modes(n, mx, my)
    x = np.arange(n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    F = np.zeros((2 * len(md), n * n))
    G = np.zeros_like(F)
    for m, (kx, ky) in enumerate(md):
        th = 2 * np.pi * ((kx * X + ky * Y) % n) / n
        c, s = np.cos(th).ravel() / n, np.sin(th).ravel() / n
        w = 1.0 if ky == 0 else 2.0
        F[2 * m], F[2 * m + 1] = c, -s
        G[2 * m], G[2 * m + 1] = w * c, -w * s
    return F, G


def pack(wre, wim):
    """(Ci, Co, Mo) complex weights -> (Mo, 2Ci, 2Co) real blocks
    [[Re, Im], [-Im, Re]]."""
    Ci, Co, Mo = wre.shape
    out = np.zeros((Mo, 2 * Ci, 2 * Co))
    for m in range(Mo):
        out[m, :Ci, :Co] = wre[:, :, m]
        out[m, :Ci, Co:] = wim[:, :, m]
        out[m, Ci:, :Co] = -wim[:, :, m]
        out[m, Ci:, Co:] = wre[:, :, m]
    return out


def forward(p, rho, cfg, emulate=True):
    """p: dict of float64 arrays (names of SpectralOperator); rho (B, n, n).
    Returns (out (B, n_out), cache)."""
    R = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    n, mx, my, L = cfg["n"], cfg["modes"][0], cfg["modes"][1], cfg["layers"]
    B = rho.shape[0]
    P = n * n
    F, G = bases(n, mx, my)
    Fr, Gr = R(F), R(G)
    ax = (np.arange(n) + 0.5) / n
    x = np.stack([rho.reshape(B, P), np.broadcast_to(np.repeat(ax, n), (B, P)),
                  np.broadcast_to(np.tile(ax, n), (B, P))], axis=1)[:, :cfg["in_ch"]]
    c = {"x": R(x), "v": [], "pre": [], "vh": [], "F": Fr, "G": Gr}
    pre = R(p["lift_w"]) @ c["x"] + p["lift_b"][None, :, None]
    v = R(pre)
    for l in range(L):
        c["v"].append(v)
        C = v.shape[1]
        vh = R(v @ Fr.T)                                      # (B, C, 2Mo)
        c["vh"].append(vh)
        Mo = vh.shape[2] // 2
        vh_m = np.concatenate([vh[:, :, 0::2], vh[:, :, 1::2]], axis=1)   # (B, 2C, Mo)  [Re; Im]
        Wp = R(pack(p[f"s{l}_re"], p[f"s{l}_im"]))           # (Mo, 2C, 2C)
        uh_m = R(np.einsum("bkm,mkn->bnm", vh_m, Wp))          # (B, 2C, Mo)
        uh = np.zeros((B, C, 2 * Mo))
        uh[:, :, 0::2], uh[:, :, 1::2] = uh_m[:, :C], uh_m[:, C:]
        u = uh @ Gr + R(p[f"w{l}"]) @ v
        u = u + p[f"b{l}"][None, :, None]
        c["pre"].append(u)
        v = R(gelu(u) if l < L - 1 else u)
    c["v"].append(v)
    hq = R(p["q_w"]) @ v + p["q_b"][None, :, None]
    c["hq"] = hq
    hm = R(gelu(hq).mean(axis=2))
    c["hm"] = hm
    out = hm @ R(p["o_w"]).T + p["o_b"][None, :]
    c["emulate"] = emulate
    return out, c


def backward(p, c, dout, cfg):
    """Gradients of sum(out * dout): (dict of parameter gradients, drho)."""
    R = bf16 if c["emulate"] else (lambda a: np.asarray(a, dtype=np.float64))
    L, n = cfg["layers"], cfg["n"]
    g = {}
    d = R(dout)
    g["o_b"] = dout.sum(axis=0)
    g["o_w"] = d.T @ c["hm"]
    dhm = d @ R(p["o_w"])
    P = c["hq"].shape[2]
    dq = dhm[:, :, None] / P * gelu_d(c["hq"])
    g["q_b"] = dq.sum(axis=(0, 2))
    dq16 = R(dq)
    g["q_w"] = (dq16 @ np.swapaxes(c["v"][L], 1, 2)).sum(axis=0)
    dv = R(p["q_w"]).T @ dq16
    Fr, Gr = c["F"], c["G"]
    for l in range(L - 1, -1, -1):
        v = c["v"][l]
        dp = dv * gelu_d(c["pre"][l]) if l < L - 1 else dv
        g[f"b{l}"] = dp.sum(axis=(0, 2))
        dp16 = R(dp)
        g[f"w{l}"] = (dp16 @ np.swapaxes(v, 1, 2)).sum(axis=0)
        duh = R(dp16 @ Gr.T)                                  # (B, C, 2Mo)
        C = v.shape[1]
        duh_m = np.concatenate([duh[:, :, 0::2], duh[:, :, 1::2]], axis=1)
        vh = c["vh"][l]
        vh_m = np.concatenate([vh[:, :, 0::2], vh[:, :, 1::2]], axis=1)
        dWp = np.einsum("bkm,bnm->mkn", vh_m, duh_m)         # (Mo, 2C, 2C)
        g[f"s{l}_re"] = np.transpose(dWp[:, :C, :C] + dWp[:, C:, C:], (1, 2, 0))
        g[f"s{l}_im"] = np.transpose(dWp[:, :C, C:] - dWp[:, C:, :C], (1, 2, 0))
        Wp = R(pack(p[f"s{l}_re"], p[f"s{l}_im"]))
        dvh_m = R(np.einsum("bnm,mkn->bkm", duh_m, Wp))
        dvh = np.zeros_like(duh)
        dvh[:, :, 0::2], dvh[:, :, 1::2] = dvh_m[:, :C], dvh_m[:, C:]
        dv = R(p[f"w{l}"]).T @ dp16 + dvh @ Fr
    g["lift_b"] = dv.sum(axis=(0, 2))
    dp16 = R(dv)
    g["lift_w"] = (dp16 @ np.swapaxes(c["x"], 1, 2)).sum(axis=0)
    drho = R(p["lift_w"])[:, 0] @ dp16
    return g, drho.reshape(-1, n, n)
