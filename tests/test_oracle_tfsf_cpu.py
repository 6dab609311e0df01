"""Pin of the oracle's TFSF magnetic currents (stepper_tm.py:169-193,
stepper_te.py:120-128: H -= (dt * h_amp) * mask in the B phase), which no
golden vector covers, by what the physics fixes.

A line current J on the E grid plus a magnetic line current M half a cell
upstream, driven with the reference's TFSF timing (ports/mode_source.py:
TM h_amplitude = -f(t dt - dx/2) with J at (t + 1/2) dt; TE h_amplitude =
+f((t + 1/2) dt - dx/2) with J at t dt), is a Huygens source: in a uniform
medium it radiates to +x only, with twice the field (four times the flux) of
J alone, and with the sign of M flipped it radiates to -x instead.  A wrong
sign, a wrong component or a wrong half-step of M fails one of these.  The
currents also superpose linearly with the electric source."""
import math

import pytest
import torch

from oracle import yee2d

torch.set_num_threads(8)
NX, NY, NPML, DX, N = 200, 80, 16, 0.05, 900


def pulse(t, fc=1.0, fw=0.6, cutoff=5.0):
    w = 1.0 / fw
    t0 = cutoff * w
    return math.cos(2 * math.pi * fc * (t - t0)) * math.exp(-((t - t0) ** 2) / (2 * w * w))


def _setup(mode):
    dt_ = torch.float64
    eps = torch.ones(NX, NY, dtype=dt_)
    if mode == "tm":
        c = yee2d.tm_coeffs(NX, NY, DX, eps, NPML, dtype=dt_)
        lo, hi = NPML, NY - NPML
        mask = torch.zeros(NX, NY, dtype=dt_)
        hm = torch.zeros(NX - 1, NY, dtype=dt_)
        key, kind = "hy_mask", "tm_x"
        amps = [pulse((t + 0.5) * c.dt) for t in range(N)]
        h_of = lambda sgn: [-sgn * pulse(t * c.dt - DX / 2) for t in range(N)]
        state = yee2d.tm_state(NX, NY, dt_)
    else:
        c = yee2d.te_coeffs(NX, NY, DX, eps, NPML, dtype=dt_)
        lo, hi = NPML, NY - 1 - NPML
        mask = torch.zeros(NX, NY - 1, dtype=dt_)
        hm = torch.zeros(NX - 1, NY - 1, dtype=dt_)
        key, kind = "hz_mask", "te_x"
        amps = [pulse(t * c.dt) for t in range(N)]
        h_of = lambda sgn: [sgn * pulse((t + 0.5) * c.dt - DX / 2) for t in range(N)]
        state = yee2d.te_state(NX, NY, dt_)
    mask[100, lo:hi] = 1.0 / DX
    hm[99, lo:hi] = 1.0 / DX
    mons = [{"type": "flux", "kind": kind, "idx": 150, "lo": lo, "hi": hi, "freqs": [0.8, 1.0, 1.2]},
            {"type": "flux", "kind": kind, "idx": 50, "lo": lo, "hi": hi, "freqs": [0.8, 1.0, 1.2]}]
    return c, state, mask, hm, key, kind, amps, h_of, mons


def _fluxes(mode, sign, e_on=True, n=N):
    c, st, mask, hm, key, kind, amps, h_of, mons = _setup(mode)
    kw = {} if sign is None else {"hamps": h_of(sign), "hmasks": {key: hm}}
    a = amps if e_on else [0.0] * N
    s, acc = yee2d.run2d(mode, st, c, n, amps=a, mask=mask, monitors=mons, **kw)
    return (yee2d.flux_from_acc(kind, acc[0], DX), yee2d.flux_from_acc(kind, acc[1], DX), s)


@pytest.mark.parametrize("mode", ["tm", "te"])
def test_huygens_pair_radiates_one_way(mode):
    f_e, b_e, _ = _fluxes(mode, None)
    f_p, b_p, _ = _fluxes(mode, +1.0)
    f_m, b_m, _ = _fluxes(mode, -1.0)
    # J alone: symmetric up to the half-cell offset of the averaged H term
    # of the two monitor lines (3e-6 relative here)
    assert torch.allclose(f_e, -b_e, rtol=1e-4, atol=0)
    # J + M: forward only, twice the field
    assert float((b_p.abs() / f_p).max()) < 2e-3
    assert abs(float(f_p[1] / f_e[1]) - 4.0) < 0.2
    # M flipped: backward only
    assert float((f_m / b_m.abs()).max()) < 0.05


@pytest.mark.parametrize("mode", ["tm", "te"])
def test_magnetic_current_superposes(mode):
    """Mid-pulse (400 steps, fields near their peak): J + M equals J plus M
    to FP64 rounding."""
    _, _, s_both = _fluxes(mode, +1.0, n=400)
    _, _, s_e = _fluxes(mode, None, n=400)
    _, _, s_h = _fluxes(mode, +1.0, e_on=False, n=400)
    for f in s_both:
        scale = max(float(s_e[f].abs().max()), float(s_h[f].abs().max()))
        assert scale > 0, f
        assert float((s_both[f] - (s_e[f] + s_h[f])).abs().max()) <= 1e-12 * scale, f
