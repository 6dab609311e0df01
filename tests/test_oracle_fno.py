"""Pins the CPU oracle of the spectral neural-operator surrogate
(oracle/fno.py) without trusting its own formulas: truncated-DFT bases
against numpy.fft, forward and gradients against an independent
torch.fft.rfft2/irfft2 + torch.autograd formulation, the backward against
central finite differences, the BF16 emulation against torch's rounding."""
import numpy as np
import pytest
import torch

from oracle import fno as O

CFG = dict(n=12, width=5, modes=(4, 3), layers=3, head=6, n_out=4, in_ch=3)


def _params(cfg, seed=0):
    rng = np.random.default_rng(seed)
    C, H, Mo = cfg["width"], cfg["head"], cfg["modes"][0] * cfg["modes"][1]
    p = {"lift_w": rng.normal(size=(C, cfg["in_ch"])), "lift_b": rng.normal(size=C) * 0.1}
    for l in range(cfg["layers"]):
        p[f"s{l}_re"] = rng.normal(size=(C, C, Mo)) * 0.3
        p[f"s{l}_im"] = rng.normal(size=(C, C, Mo)) * 0.3
        p[f"w{l}"] = rng.normal(size=(C, C)) * 0.4
        p[f"b{l}"] = rng.normal(size=C) * 0.1
    p.update(q_w=rng.normal(size=(H, C)) * 0.4, q_b=rng.normal(size=H) * 0.1,
             o_w=rng.normal(size=(cfg["n_out"], H)) * 0.4, o_b=rng.normal(size=cfg["n_out"]) * 0.1)
    return p


def _torch_net(p, rho, cfg):
    """Same architecture written with torch.fft (not the oracle's DFT bases)."""
    n, (mx, my), L = cfg["n"], cfg["modes"], cfg["layers"]
    B = rho.shape[0]
    ax = (torch.arange(n, dtype=torch.float64) + 0.5) / n
    x = torch.stack([rho, ax.view(n, 1).expand(B, n, n), ax.view(1, n).expand(B, n, n)], 1)[:, :cfg["in_ch"]]
    gelu = torch.nn.functional.gelu
    v = torch.einsum("oc,bcxy->boxy", p["lift_w"], x) + p["lift_b"].view(1, -1, 1, 1)
    kxs = [a if a < mx // 2 else n - mx + a for a in range(mx)]
    for l in range(L):
        C = v.shape[1]
        Vh = torch.fft.rfft2(v, norm="ortho")[:, :, kxs, :my]
        W = torch.complex(p[f"s{l}_re"], p[f"s{l}_im"]).view(C, C, mx, my)
        Us = torch.einsum("bixy,ioxy->boxy", Vh, W)
        Z = torch.zeros((B, C, n, n // 2 + 1), dtype=torch.complex128)
        Z[:, :, kxs, :my] = Us
        u = torch.fft.irfft2(Z, s=(n, n), norm="ortho") + torch.einsum("oc,bcxy->boxy", p[f"w{l}"], v)
        u = u + p[f"b{l}"].view(1, -1, 1, 1)
        v = gelu(u) if l < L - 1 else u
    h = gelu(torch.einsum("hc,bcxy->bhxy", p["q_w"], v) + p["q_b"].view(1, -1, 1, 1))
    return h.mean((2, 3)) @ p["o_w"].T + p["o_b"]


def test_bases_match_numpy_fft():
    n, mx, my = 16, 6, 4
    F, G = O.bases(n, mx, my)
    rng = np.random.default_rng(1)
    v = rng.normal(size=(n, n))
    spec = np.fft.rfft2(v, norm="ortho")
    md = O.modes(n, mx, my)
    got = F @ v.ravel()
    want = np.array([[spec[kx, ky].real, spec[kx, ky].imag] for kx, ky in md]).ravel()
    assert np.abs(got - want).max() < 1e-12
    U = rng.normal(size=len(md)) + 1j * rng.normal(size=len(md))
    Z = np.zeros((n, n // 2 + 1), complex)
    for (kx, ky), u in zip(md, U):
        Z[kx, ky] = u
    want = np.fft.irfft2(Z, s=(n, n), norm="ortho").ravel()
    got = np.stack([U.real, U.imag], 1).ravel() @ G
    assert np.abs(got - want).max() < 1e-12


def test_bf16_emulation_matches_torch_rounding():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(size=5000) * 10.0 ** rng.integers(-8, 8, 5000),
                        (np.arange(-300, 300) * 2.0 ** -10 + 1.0 / 512)])  # exact ties
    want = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).double().numpy()
    assert np.array_equal(O.bf16(x), want)


def test_forward_and_gradients_match_torch_autograd():
    cfg = CFG
    p = _params(cfg)
    rng = np.random.default_rng(3)
    rho = rng.uniform(size=(2, cfg["n"], cfg["n"]))
    dout = rng.normal(size=(2, cfg["n_out"]))
    out, cache = O.forward(p, rho, cfg, emulate=False)
    g, drho = O.backward(p, cache, dout, cfg)
    tp = {k: torch.tensor(v, requires_grad=True) for k, v in p.items()}
    trho = torch.tensor(rho, requires_grad=True)
    tout = _torch_net(tp, trho, cfg)
    assert np.abs(tout.detach().numpy() - out).max() < 1e-12 * max(1.0, np.abs(out).max())
    (tout * torch.tensor(dout)).sum().backward()
    for k in p:
        ref = tp[k].grad.numpy()
        assert np.abs(g[k] - ref).max() <= 1e-11 * max(1.0, np.abs(ref).max()), k
    assert np.abs(drho - trho.grad.numpy()).max() <= 1e-11 * max(1.0, np.abs(drho).max())


def test_backward_central_fd():
    cfg = dict(CFG, layers=2)
    p = _params(cfg, seed=4)
    rng = np.random.default_rng(5)
    rho = rng.uniform(size=(1, cfg["n"], cfg["n"]))
    dout = rng.normal(size=(1, cfg["n_out"]))
    out, cache = O.forward(p, rho, cfg, emulate=False)
    g, drho = O.backward(p, cache, dout, cfg)

    def J(pp, rr):
        return float((O.forward(pp, rr, cfg, emulate=False)[0] * dout).sum())

    h = 1e-6
    for name, idx in [("s1_re", (1, 2, 5)), ("s0_im", (0, 4, 11)), ("w0", (2, 3)), ("lift_w", (4, 0)),
                      ("q_w", (5, 1)), ("b1", (2,)), ("o_b", (3,))]:
        pp, pm = {k: v.copy() for k, v in p.items()}, {k: v.copy() for k, v in p.items()}
        pp[name][idx] += h
        pm[name][idx] -= h
        fd = (J(pp, rho) - J(pm, rho)) / (2 * h)
        assert abs(g[name][idx] - fd) <= 1e-6 * max(1.0, abs(fd)), (name, g[name][idx], fd)
    for (i, j) in [(3, 4), (7, 1), (11, 10)]:
        rp, rm = rho.copy(), rho.copy()
        rp[0, i, j] += h
        rm[0, i, j] -= h
        fd = (J(p, rp) - J(p, rm)) / (2 * h)
        assert abs(drho[0, i, j] - fd) <= 1e-6 * max(1.0, abs(fd))


def test_gelu_derivative_fd():
    x = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (O.gelu(x + h) - O.gelu(x - h)) / (2 * h)
    assert np.abs(O.gelu_d(x) - fd).max() < 1e-8
    assert np.allclose(O.gelu(x), torch.nn.functional.gelu(torch.tensor(x)).numpy(), atol=1e-15)


def test_pack_is_complex_multiplication():
    rng = np.random.default_rng(6)
    wre, wim = rng.normal(size=(3, 4, 5)), rng.normal(size=(3, 4, 5))
    V = rng.normal(size=(5, 3)) + 1j * rng.normal(size=(5, 3))
    Wp = O.pack(wre, wim)
    for m in range(5):
        got = np.concatenate([V[m].real, V[m].imag]) @ Wp[m]
        want = V[m] @ (wre[:, :, m] + 1j * wim[:, :, m])
        assert np.allclose(got, np.concatenate([want.real, want.imag]), atol=1e-13)
