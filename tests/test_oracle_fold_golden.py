"""Pin the oracle's symmetry folds to the reference bit for bit: the fold_2d
golden vectors (tests/golden/make_golden.py) hold the reference's own runs of
2D TM cells cut by an even / odd y mirror, an even / odd x mirror, both
mirrors, Rotate2 and Rotate4 (each phase), and TE cells cut by even / odd y
and x mirrors, in FP32 and FP64 (meep_torch/core/coeffs.py:211-281,457-469,
stepper_tm.py:94-155, stepper_te.py:80-97).  Every field, the on-plane H rows
Hx0 / Hy0 and the flux accumulator must match exactly: a wrong ghost sign, a
transposed Rotate4 ghost, an unflipped Rotate2 ghost or a missing on-plane
row fails here."""
import numpy as np
import pytest
import torch

from conftest import golden
from oracle import yee2d

N, STEPS = 36, 200
FOLD_TM = {"ym+": dict(y_mirror=1), "ym-": dict(y_mirror=-1), "xm+": dict(x_mirror=1),
           "xm-": dict(x_mirror=-1), "xy": dict(x_mirror=-1, y_mirror=1), "r2+": dict(y_rot2=1),
           "r2-": dict(y_rot2=-1), "r4+": dict(rot4=1), "r4-": dict(rot4=-1)}
FOLD_TE = {"ym+": dict(y_mirror=1), "ym-": dict(y_mirror=-1), "xm+": dict(x_mirror=1),
           "xm-": dict(x_mirror=-1)}
FREQS = [0.7, 0.9, 1.1]


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a))


def tm_npml(kw):
    return ((0 if kw.get("x_mirror") or kw.get("rot4") else 6, 6),
            (0 if kw.get("y_mirror") or kw.get("y_rot2") or kw.get("rot4") else 6, 6))


def te_npml(kw):
    return ((0 if kw.get("x_mirror") else 6, 6), (0 if kw.get("y_mirror") else 6, 6))


@pytest.mark.parametrize("name", sorted(FOLD_TM))
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_tm_folds_bitexact(name, tag, dtype):
    g = golden(f"fold_2d_{tag}")
    kw = FOLD_TM[name]
    c = yee2d.tm_coeffs(N, N, 0.05, _t(g["eps"]), tm_npml(kw), dtype=dtype, **kw)
    mons = [{"type": "flux", "kind": "tm_x", "idx": 26, "lo": 0, "hi": 28, "freqs": FREQS}]
    s, acc = yee2d.run2d("tm", yee2d.tm_state(N, N, dtype, **kw), c, STEPS, amps=g["amp"],
                         mask=_t(g[f"tm{name}_mask"]), monitors=mons)
    keys = [k[len(f"tm{name}_"):] for k in g.keys() if k.startswith(f"tm{name}_") and not k.endswith(("_mask", "_acc"))]
    assert {"Ez", "Hx", "psi_Hxy"} <= set(keys)
    for f in keys:
        assert torch.equal(s[f], _t(g[f"tm{name}_{f}"])), (name, f)
    assert torch.equal(acc[0], _t(g[f"tm{name}_acc"]))


@pytest.mark.parametrize("name", sorted(FOLD_TE))
@pytest.mark.parametrize("tag,dtype", [("f32", torch.float32), ("f64", torch.float64)])
def test_te_folds_bitexact(name, tag, dtype):
    g = golden(f"fold_2d_{tag}")
    kw = FOLD_TE[name]
    c = yee2d.te_coeffs(N, N, 0.05, _t(g["eps"]), te_npml(kw), dtype=dtype, **kw)
    mons = [{"type": "flux", "kind": "te_x", "idx": 26, "lo": 0, "hi": 28, "freqs": FREQS}]
    s, acc = yee2d.run2d("te", yee2d.te_state(N, N, dtype), c, STEPS, amps=g["amp"],
                         mask=_t(g[f"te{name}_mask"]), monitors=mons)
    for f in ("Hz", "Ex", "Ey", "Dx", "Dy", "psi_Eyx", "psi_Exy", "psi_Hzy", "psi_Hzx"):
        assert torch.equal(s[f], _t(g[f"te{name}_{f}"])), (name, f)
    assert torch.equal(acc[0], _t(g[f"te{name}_acc"]))


def test_parities_differ():
    """The fixtures pin the parity, not only the geometry: the even and odd
    y-mirror runs of the reference differ by far more than rounding, and so
    do the two Rotate4 phases."""
    g = golden("fold_2d_f64")
    for a, b in (("tmym+", "tmym-"), ("tmr4+", "tmr4-"), ("teym+", "teym-")):
        f = "Ez" if a.startswith("tm") else "Hz"
        x, y = g[f"{a}_{f}"], g[f"{b}_{f}"]
        assert np.linalg.norm(x - y) > 1e-3 * np.linalg.norm(x)
