"""Pin of the oracle's TM media terms, which no golden vector covers, by the
physics the reference's own analytic tests use:

* electric ADE poles (coeffs.py:87-106, stepper_tm.py:203-231): a 2 um
  Lorentz slab (eps_inf 2.25, pole fn 2, gamma 0.1, sigma 1.5) against the
  transfer-matrix transmission with eps(f) = eps_inf + sigma fn^2 /
  (fn^2 - f^2 - i f gamma), within 3% (the scene and bar of
  meep_torch tests/analytic/test_dispersion.py);
* sigma_d (coeffs.py:161-163, stepper_tm.py:161-163, 198-200): an L = 2 slab
  with sigma_d = 1 against the analytic amplitude transmission
  exp(-alpha L) |4 n / (1 + n)^2|, n = sqrt(1 + i sigma / w), within 2%, and
  against meep's own transmission for the same slab (the reference's fixture
  tests/examples/data/conductive_slab.json, copied to
  tests/golden/meep_conductive_slab.json), amplitude rms within 3% (the bars
  of meep_torch tests/examples/test_w5_conductivity.py);
* chi3 (stepper_tm.py:228-231): third-harmonic generation scales as the cube
  of the drive (2x input -> 8x, within 5%) and vanishes without chi3
  (meep_torch tests/analytic/test_kerr.py);
* the TE step's per-axis poles and sigma_d on the same slabs, as a TE
  parallel-plate wave (PEC y walls, where the y-uniform Ey / Hz wave is
  exact).

A dropped pole term, a wrong pole coefficient, a wrong conductivity factor or
a missing cubic term moves these by far more than the bars."""
import cmath
import json
import math
import os

import numpy as np
import pytest
import torch

from oracle import yee2d

torch.set_num_threads(8)
DT = torch.float64
HERE = os.path.dirname(os.path.abspath(__file__))


def gaussian(fc, fw, cutoff=5.0):
    w = 1.0 / fw
    t0 = cutoff * w
    return lambda t: math.cos(2 * math.pi * fc * (t - t0)) * math.exp(-((t - t0) ** 2) / (2 * w * w))


def _line_run(nx, dx, c, src_ix, amps, mons, n_poles=0):
    ny = 8
    mask = torch.zeros(nx, ny, dtype=DT)
    mask[src_ix, :] = 1.0 / dx
    st = yee2d.tm_state(nx, ny, DT, y_pmc=True, n_poles=n_poles)
    return yee2d.run2d("tm", st, c, len(amps), amps=amps, mask=mask, monitors=mons)


def test_lorentz_slab_matches_transfer_matrix():
    nx, ny, dx = 1400, 8, 0.02
    pole, eps_inf = (2.0, 0.1, 1.5), 2.25
    freqs = [0.50, 0.55, 0.60, 0.65, 0.70, 0.75, 0.80]

    def flux(slab_on):
        slab = torch.zeros(nx, ny, dtype=DT)
        slab[600:700, :] = 1.0                        # 2 um
        eps = 1.0 + (eps_inf - 1.0) * slab if slab_on else torch.ones(nx, ny, dtype=DT)
        poles = [pole] if slab_on else []
        c = yee2d.tm_coeffs(nx, ny, dx, eps, 40, dtype=DT, y_pmc=True, poles=poles,
                            pole_masks=[slab] if slab_on else None)
        f = gaussian(0.65, 0.35)
        amps = [f((t + 0.5) * c.dt) for t in range(9000)]
        mons = [{"type": "flux", "kind": "tm_x", "idx": 1200, "lo": 0, "hi": ny, "freqs": freqs}]
        _, acc = _line_run(nx, dx, c, 100, amps, mons, n_poles=len(poles))
        return yee2d.flux_from_acc("tm_x", acc[0], dx)

    T = flux(True) / flux(False)
    for i, f in enumerate(freqs):
        fn, g, s = pole
        n2 = cmath.sqrt(eps_inf + s * fn * fn / (fn * fn - f * f - 1j * f * g))
        delta = 2 * cmath.pi * f * n2 * 2.0
        r12 = (1 - n2) / (1 + n2)
        t = (2 / (1 + n2)) * (2 * n2 / (1 + n2)) * cmath.exp(1j * delta) / (1 - r12 ** 2 * cmath.exp(2j * delta))
        assert abs(float(T[i]) - abs(t) ** 2) / abs(t) ** 2 < 0.03, (f, float(T[i]), abs(t) ** 2)


def _conductive_T(freqs):
    """Power transmission through x in [14, 16] um (sigma_d = 1) of a 30 um
    line at resolution 40, 1.5 um PML, source at 5 um, flux at 21 um."""
    dx = 1.0 / 40
    nx, ny = int(round(30 / dx)), 8

    def flux(on):
        slab = torch.zeros(nx, ny, dtype=DT)
        slab[int(round(14 / dx)):int(round(16 / dx)), :] = 1.0
        c = yee2d.tm_coeffs(nx, ny, dx, torch.ones(nx, ny, dtype=DT), int(round(1.5 / dx)), dtype=DT,
                            y_pmc=True, sigma_d=slab if on else None)
        f = gaussian(1.0, 0.2)
        amps = [f((t + 0.5) * c.dt) for t in range(10000)]
        mons = [{"type": "flux", "kind": "tm_x", "idx": int(round(21 / dx)), "lo": 0, "hi": ny, "freqs": freqs}]
        _, acc = _line_run(nx, dx, c, int(round(5 / dx)), amps, mons)
        return yee2d.flux_from_acc("tm_x", acc[0], dx)

    return (flux(True) / flux(False)).numpy()


def test_conductive_slab_matches_analytic_and_meep():
    with open(os.path.join(HERE, "golden", "meep_conductive_slab.json")) as fh:
        ref = json.load(fh)
    T = _conductive_T(ref["freqs"])
    w = 2 * math.pi * 1.0
    n = complex(np.sqrt(1.0 + 1j * ref["sigma_d"] / w))
    t_an = math.exp(-w * abs(n.imag) * ref["L"]) * abs(4 * n / (1 + n) ** 2)
    amp = np.sqrt(T)
    assert abs(amp[2] - t_an) / t_an < 0.02, (amp[2], t_an)          # f = fcen
    rms = float(np.sqrt(np.mean((amp / np.sqrt(ref["T_power"]) - 1.0) ** 2)))
    assert rms < 0.03, rms


def test_kerr_third_harmonic_scales_as_cube():
    f0 = 0.25

    def peaks(amp, chi3):
        nx, ny, dx = 1000, 8, 0.02
        c = yee2d.tm_coeffs(nx, ny, dx, torch.full((nx, ny), 2.25, dtype=DT), 40, dtype=DT, y_pmc=True,
                            chi3=chi3)

        def a(t):
            ramp = 1.0 if t >= 40.0 else 0.5 * (1 - math.cos(math.pi * t / 40.0))
            return amp * ramp * math.sin(2 * math.pi * f0 * t)

        amps = [a((t + 0.5) * c.dt) for t in range(16000)]
        _, acc = _line_run(nx, dx, c, 100, amps, [{"type": "probe", "ix": 700, "iy": 4}])
        steady = acc[0][10000:]
        spec = torch.fft.rfft(steady).abs()
        fr = torch.fft.rfftfreq(len(steady), c.dt)

        def pk(f):
            i = int((fr - f).abs().argmin())
            return float(spec[i - 2:i + 3].max())
        return pk(f0), pk(3 * f0)

    fa, ta = peaks(0.02, 0.02)
    fb, tb = peaks(0.04, 0.02)
    f0_, t0_ = peaks(0.02, 0.0)
    assert abs(fb / fa - 2.0) < 0.01
    assert abs(tb / ta - 8.0) < 0.4
    assert t0_ / f0_ < 1e-9 < ta / fa


def _te_line(nx, ny, dx, c, src_ix, amps, mons, np_=0):
    """TE parallel plate: PEC y walls, PML along x only, a line current on
    every Ey row: the y-uniform Ey / Hz plane wave is exact (Ey is normal to
    the walls), so 1D transmission formulas apply."""
    mask = torch.zeros(nx, ny - 1, dtype=DT)
    mask[src_ix, :] = 1.0 / dx
    st = yee2d.te_state(nx, ny, DT, npx=np_, npy=np_)
    return yee2d.run2d("te", st, c, len(amps), amps=amps, mask=mask, monitors=mons)


def test_te_lorentz_slab_matches_transfer_matrix():
    """The TE step's per-axis poles (stepper_te.py:137-147) on the Lorentz
    slab of the TM test, as a parallel-plate TE wave: same 3% bar."""
    nx, ny, dx = 1400, 8, 0.02
    pole, eps_inf = (2.0, 0.1, 1.5), 2.25
    freqs = [0.50, 0.60, 0.70, 0.80]

    def flux(on):
        slab = torch.zeros(nx, ny, dtype=DT)
        slab[600:700, :] = 1.0
        eps = 1.0 + (eps_inf - 1.0) * slab if on else torch.ones(nx, ny, dtype=DT)
        poles = [pole] if on else []
        c = yee2d.te_coeffs(nx, ny, dx, eps, ((40, 40), (0, 0)), dtype=DT, poles=poles,
                            pole_masks=[slab] if on else None)
        f = gaussian(0.65, 0.35)
        amps = [f((t + 0.5) * c.dt) for t in range(9000)]
        mons = [{"type": "flux", "kind": "te_x", "idx": 1200, "lo": 0, "hi": ny - 1, "freqs": freqs}]
        _, acc = _te_line(nx, ny, dx, c, 100, amps, mons, np_=len(poles))
        return yee2d.flux_from_acc("te_x", acc[0], dx)

    T = flux(True) / flux(False)
    for i, f in enumerate(freqs):
        fn, g, s = pole
        n2 = cmath.sqrt(eps_inf + s * fn * fn / (fn * fn - f * f - 1j * f * g))
        delta = 2 * cmath.pi * f * n2 * 2.0
        r12 = (1 - n2) / (1 + n2)
        t = (2 / (1 + n2)) * (2 * n2 / (1 + n2)) * cmath.exp(1j * delta) / (1 - r12 ** 2 * cmath.exp(2j * delta))
        assert abs(float(T[i]) - abs(t) ** 2) / abs(t) ** 2 < 0.03, (f, float(T[i]), abs(t) ** 2)


def test_te_conductive_slab_matches_analytic():
    """TE sigma_d (coeffs.py:399-404, stepper_te.py:96-99, 131-134) on the
    sigma_d = 1, L = 2 slab: amplitude transmission within 2% of
    exp(-alpha L) |4 n / (1 + n)^2| at f = 1."""
    dx = 1.0 / 40
    nx, ny = int(round(30 / dx)), 8

    def flux(on):
        slab = torch.zeros(nx, ny, dtype=DT)
        slab[int(round(14 / dx)):int(round(16 / dx)), :] = 1.0
        c = yee2d.te_coeffs(nx, ny, dx, torch.ones(nx, ny, dtype=DT), ((int(round(1.5 / dx)),) * 2, (0, 0)),
                            dtype=DT, sigma_d=slab if on else None)
        f = gaussian(1.0, 0.2)
        amps = [f((t + 0.5) * c.dt) for t in range(10000)]
        mons = [{"type": "flux", "kind": "te_x", "idx": int(round(21 / dx)), "lo": 0, "hi": ny - 1,
                 "freqs": [1.0]}]
        _, acc = _te_line(nx, ny, dx, c, int(round(5 / dx)), amps, mons)
        return yee2d.flux_from_acc("te_x", acc[0], dx)

    amp = math.sqrt(float(flux(True)[0] / flux(False)[0]))
    w = 2 * math.pi
    n = complex(np.sqrt(1.0 + 1j / w))
    t_an = math.exp(-w * abs(n.imag) * 2.0) * abs(4 * n / (1 + n) ** 2)
    assert abs(amp - t_an) / t_an < 0.02, (amp, t_an)


def test_te_kerr_third_harmonic_scales_as_cube():
    """The TE chi3 fixed point (stepper_te.py:161-175) on the parallel-plate
    wave: twice the drive gives eight times the third harmonic, and none
    without chi3."""
    f0 = 0.25

    def peaks(amp, chi3):
        nx, ny, dx = 1000, 8, 0.02
        c = yee2d.te_coeffs(nx, ny, dx, torch.full((nx, ny), 2.25, dtype=DT), ((40, 40), (0, 0)), dtype=DT,
                            chi3=chi3)

        def a(t):
            ramp = 1.0 if t >= 40.0 else 0.5 * (1 - math.cos(math.pi * t / 40.0))
            return amp * ramp * math.sin(2 * math.pi * f0 * t)

        mask = torch.zeros(nx, ny - 1, dtype=DT)
        mask[100, :] = 1.0 / dx
        st, series = yee2d.te_state(nx, ny, DT), []
        for t in range(16000):
            st = yee2d.te_step(st, c, c.dt * a((t + 0.5) * c.dt), mask)
            series.append(float(st["Ey"][700, 3]))
        steady = torch.tensor(series[10000:], dtype=DT)   # past the ramp and the transit
        spec = torch.fft.rfft(steady).abs()
        fr = torch.fft.rfftfreq(len(steady), c.dt)

        def pk(f):
            i = int((fr - f).abs().argmin())
            return float(spec[i - 2:i + 3].max())
        return pk(f0), pk(3 * f0)

    fa, ta = peaks(0.02, 0.02)
    fb, tb = peaks(0.04, 0.02)
    f0_, t0_ = peaks(0.02, 0.0)
    assert abs(fb / fa - 2.0) < 0.01
    assert abs(tb / ta - 8.0) < 0.4
    assert t0_ / f0_ < 1e-9 < ta / fa
