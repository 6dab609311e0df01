"""Mode amplitudes from DFT line fields (meep_torch/ports/mode_monitor.py:1-68).

Bi-orthogonal overlap <E1 x H2> without conjugation.  For mode m with
transverse profiles (Em, Hm) on the port line and the DFT'd line fields
(E, H) at one frequency:

    N = sum(Em Hm) dx,  EH = sum(E Hm) dx,  HE = sum(Em H) dx,
    a+ = (EH + HE) / (2 N) sqrt|N|,  a- = (EH - HE) / (2 N) sqrt|N|,

so that |a|^2 scales with modal power.  TM: Em = profile,
Hm = axis_sign neff Em (axis_sign -1 for x-normal lines, whose +x wave
stores Hy = -neff Ez; +1 for y-normal).  TE (x-normal): Hm = profile,
Em = neff Hm / eps."""
import torch

from .modes1d import solve_1d

NORM_FLOOR = 1e-12


def dft_fields(acc):
    """(nfreq, 4, L) flux accumulator [Re E, Im E, Re H, Im H] -> complex E, H."""
    re_e, im_e, re_h, im_h = acc.unbind(1)
    return torch.complex(re_e, im_e), torch.complex(re_h, im_h)


def decompose(acc, eps_line, dx, freqs, pol, n_modes, axis_sign=-1.0, bc="pec"):
    """Forward / backward amplitudes (a_p, a_m), each (nfreq, n_modes)
    complex128 on the accumulator's device; one mode solve per frequency."""
    E, H = dft_fields(acc)
    dev = E.device
    nf = len(freqs)
    a_p = torch.zeros(nf, n_modes, dtype=torch.complex128, device=dev)
    a_m = torch.zeros_like(a_p)
    eps_line = eps_line.detach()
    for i, f in enumerate(freqs):
        modes = solve_1d(eps_line, dx, f, pol, k=n_modes, bc=bc)
        for m in range(n_modes):
            prof = modes.profiles[m].to(dev)
            neff = float(modes.neff[m])
            if pol == "tm":
                em, hm = prof, axis_sign * neff * prof
            else:
                hm = prof
                em = neff * prof / eps_line.to(device=dev, dtype=prof.dtype)
            norm = float((em * hm).sum()) * dx
            if abs(norm) < NORM_FLOOR:
                raise ValueError(f"degenerate/leaky mode {m} at f={f}: norm {norm:.3e}")
            eh = (E[i] * hm.to(torch.complex128)).sum() * dx
            he = (em.to(torch.complex128) * H[i]).sum() * dx
            scale = abs(norm) ** 0.5 / (2.0 * norm)
            a_p[i, m] = (eh + he) * scale
            a_m[i, m] = (eh - he) * scale
    return a_p, a_m
