"""Host rules of the facade symmetries (meep_torch/mp/simulation.py:383-494,
874-970): which symmetry lists fold the cell and which raise, the index
folding of lines and points, and the full-cell reconstruction of a kept
part, checked against the closed forms of a mirror and a rotation."""
import numpy as np
import pytest


class _GS:
    def __init__(self, nx, ny):
        self.nx, self.ny = nx, ny


def _fold(syms, nx=40, ny=40, pol="tm", npml=((5, 5), (5, 5)), nabs=((0, 0), (0, 0))):
    from meep_torch_amd.mp.symmetry import Folding
    return Folding(syms, _GS(nx, ny), pol, [list(p) for p in npml], [list(p) for p in nabs])


def test_cut_state_and_builder_arguments():
    import meep_torch_amd.mp as mp
    f = _fold([mp.Mirror(mp.Y, phase=-1), mp.Mirror(mp.X)])
    assert (f.cx, f.cy, f.px, f.py) == (20, 20, 1, -1)
    assert f.cut_kwargs("tm") == dict(y_mirror=-1, x_mirror=1, y_rot2=0, rot4=0)
    f = _fold([mp.Rotate2(mp.Z, phase=-1)])
    assert (f.cx, f.cy, f.rot2) == (0, 20, -1) and f.cut_kwargs("tm") == dict(y_mirror=0, x_mirror=0, y_rot2=-1,
                                                                              rot4=0)
    f = _fold([mp.Rotate4(mp.Z)])
    assert (f.cx, f.cy, f.rot4) == (20, 20, 1) and f.cut_kwargs("tm")["rot4"] == 1
    assert not _fold([mp.Identity()]).active


@pytest.mark.parametrize("syms,kw,match", [
    ("mirror_z", {}, "not supported"),
    ("rot_mirror", {}, "combined with Mirror"),
    ("r2_r4", {}, "C4 already contains C2"),
    ("odd_y", dict(ny=41), "even cell row count"),
    ("r2_pml", dict(npml=((5, 3), (5, 5))), "x-symmetric absorber"),
    ("r4_rect", dict(nx=40, ny=44), "square cell"),
])
def test_unsupported_symmetries_raise(syms, kw, match):
    import meep_torch_amd.mp as mp
    s = {"mirror_z": [mp.Mirror(mp.Z)], "rot_mirror": [mp.Rotate2(mp.Z), mp.Mirror(mp.X)],
         "r2_r4": [mp.Rotate2(mp.Z), mp.Rotate4(mp.Z)], "odd_y": [mp.Mirror(mp.Y)],
         "r2_pml": [mp.Rotate2(mp.Z)], "r4_rect": [mp.Rotate4(mp.Z)]}[syms]
    with pytest.raises(NotImplementedError, match=match):
        _fold(s, **kw)


def test_te_rotation_keeps_the_full_cell():
    import meep_torch_amd.mp as mp
    with pytest.warns(UserWarning, match="full cell"):
        f = _fold([mp.Rotate4(mp.Z)], pol="te")
    assert not f.active


def test_index_folding():
    import meep_torch_amd.mp as mp
    f = _fold([mp.Mirror(mp.Y)])
    assert f.foldpt(25, "y") == (5, False) and f.foldpt(14, "y") == (5, True)
    assert f.fold(22, 30, "y") == (2, 10, False, False)
    assert f.fold(10, 18, "y") == (2, 10, False, True)
    assert f.fold(15, 25, "y") == (0, 5, True, False)
    assert f.foldpt(3, "x") == (3, False)        # no cut on x


def test_reconstruction_closed_forms():
    """A mirror-even / odd and a C2 / C4 symmetric field, cut and rebuilt,
    is itself: the rebuild is the symmetry operation, with its phase."""
    import meep_torch_amd.mp as mp
    rng = np.random.default_rng(3)
    n = 8
    a = rng.standard_normal((n, n // 2))
    full_even = np.concatenate([a[:, ::-1], a], axis=1)
    f = _fold([mp.Mirror(mp.Y)], nx=n, ny=n)
    assert np.array_equal(f.demirror(full_even[:, n // 2:], f.parity("Ez")), full_even)
    f = _fold([mp.Mirror(mp.Y, phase=-1)], nx=n, ny=n)
    full_odd = np.concatenate([-a[:, ::-1], a], axis=1)
    assert np.array_equal(f.demirror(full_odd[:, n // 2:], f.parity("Ez")), full_odd)
    # C2: E(nx-1-i, ny-1-j) = p E(i, j)
    f = _fold([mp.Rotate2(mp.Z, phase=-1)], nx=n, ny=n, npml=((0, 0), (0, 0)))
    full = f.demirror(a, f.parity("Ez"))
    assert np.array_equal(full[::-1, ::-1], -full)
    # C4: a quarter turn maps the field onto itself times the phase
    q = rng.standard_normal((n // 2, n // 2))
    f = _fold([mp.Rotate4(mp.Z, phase=-1)], nx=n, ny=n, npml=((0, 0), (0, 0)))
    full = f.demirror(q)
    assert np.array_equal(np.rot90(full), -full) or np.array_equal(np.rot90(full, -1), -full)
    assert np.array_equal(full[n // 2:, n // 2:], q)
