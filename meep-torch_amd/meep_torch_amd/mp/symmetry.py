"""meep symmetries for 2D cells (meep_torch/mp/simulation.py:54-85 classes,
:383-494 fold rules, :874-970 index folding and output reconstruction).

A symmetry halves (Mirror, Rotate2) or quarters (Rotate4) the engine's
domain: the run keeps the cells at and above the cut plane(s) (full-cell
rows j >= cy, columns i >= cx) and the core stepper closes the cut with the
matching wall (build_tm_coeffs y_mirror / x_mirror / y_rot2 / rot4,
build_te_coeffs y_mirror / x_mirror).  Sources below a cut are dropped (the
wall supplies their images), monitor lines are folded onto the kept part,
and field arrays are rebuilt on the full cell from the kept part and the
symmetry's phase."""
import warnings

import numpy as np

from .vec import X, Y, Z


class Symmetry:
    """Symmetry of the structure and the sources; ``direction`` is the mirror
    plane's normal or the rotation axis, ``phase`` multiplies the fields
    under the operation (-1: odd)."""

    def __init__(self, direction=None, phase=1.0 + 0j):
        self.direction = direction
        self.phase = complex(phase)


class Mirror(Symmetry):
    """Mirror plane through the cell centre with normal X or Y."""


class Rotate2(Symmetry):
    """180-degree rotation about Z (C2)."""


class Rotate4(Symmetry):
    """90-degree rotation about Z (C4), square cells."""


class Identity(Symmetry):
    """No operation."""


_HINT = "Mirror(X/Y, phase=+-1), Rotate2(Z, phase=+-1) and Rotate4(Z, phase=+-1) in 2D cells"


def check_symmetry(s):
    """A symmetry the facade can fold: an instance of the classes above."""
    if not isinstance(s, Symmetry):
        raise NotImplementedError(f"symmetry {s!r} is not supported: {_HINT}")


class Folding:
    """The cut state of one run: cx / cy = full-cell columns / rows below the
    cut (0 = no cut on that axis), px / py = the mirror phases, rot2 / rot4 =
    the rotation phase (0 = none)."""

    def __init__(self, symmetries, gs, pol, npml_sides, nabs_sides):
        self.cx = self.cy = 0
        self.px = self.py = 1
        self.rot2 = self.rot4 = 0
        mirrors, rot2, rot4 = {}, 0, 0
        for s in symmetries or ():
            check_symmetry(s)
            real = s.phase in (1 + 0j, -1 + 0j)
            if isinstance(s, Identity):
                continue
            if isinstance(s, Rotate4) and s.direction == Z and real:
                rot4 = int(s.phase.real)
            elif isinstance(s, Rotate2) and s.direction == Z and real:
                rot2 = int(s.phase.real)
            elif isinstance(s, Mirror) and s.direction in (X, Y) and real:
                mirrors[s.direction] = int(s.phase.real)
            else:
                raise NotImplementedError(f"{type(s).__name__}(direction={s.direction}, phase={s.phase}) "
                                          f"is not supported: {_HINT}")
        if (rot2 or rot4) and mirrors:
            raise NotImplementedError("Rotate2/Rotate4 combined with Mirror is not supported: a rotation fold "
                                      "and a mirror cut do not compose in the half-domain engine")
        if rot2 and rot4:
            raise NotImplementedError("Rotate2 with Rotate4: C4 already contains C2, list Rotate4 alone")
        if (rot2 or rot4) and pol != "tm":
            warnings.warn("Rotate2/Rotate4 on a TE run keep the full cell")
            rot2 = rot4 = 0
        if Y in mirrors:
            if gs.ny % 2:
                raise NotImplementedError("Mirror(Y) needs an even cell row count")
            self.cy, self.py = gs.ny // 2, mirrors[Y]
        if X in mirrors:
            if gs.nx % 2:
                raise NotImplementedError("Mirror(X) needs an even cell column count")
            self.cx, self.px = gs.nx // 2, mirrors[X]
        if rot2:
            # the C2 fold maps (i, j) onto (nx-1-i, ny-1-j): the kept half is
            # the upper rows with the whole x extent, which must be
            # reversal-symmetric
            if gs.ny % 2 or gs.nx % 2:
                raise NotImplementedError("Rotate2(Z) needs even cell counts")
            if npml_sides[0][0] != npml_sides[0][1] or nabs_sides[0][0] != nabs_sides[0][1]:
                raise NotImplementedError("Rotate2(Z) needs an x-symmetric absorber: the C2 fold maps the low-x "
                                          "edge onto the high-x edge")
            self.cy, self.py, self.rot2 = gs.ny // 2, rot2, rot2
        if rot4:
            if gs.nx % 2 or gs.ny % 2 or gs.nx != gs.ny:
                raise NotImplementedError(f"Rotate4(Z) needs a square cell of even size, got {gs.nx}x{gs.ny}")
            four = lambda t: set(tuple(t[0]) + tuple(t[1]))
            if len(four(npml_sides)) > 1 or len(four(nabs_sides)) > 1:
                raise NotImplementedError("Rotate4(Z) needs the same absorber on all four edges")
            self.cx = self.cy = gs.nx // 2
            self.px = self.py = self.rot4 = rot4

    @property
    def active(self):
        return bool(self.cx or self.cy)

    def cut_kwargs(self, pol):
        """The core builders' cut arguments (mp/cpp/core.py:213-217)."""
        rot = bool(self.rot2 or self.rot4)
        kw = dict(y_mirror=self.py if (self.cy and not rot) else 0,
                  x_mirror=self.px if (self.cx and not rot) else 0)
        if pol == "tm":
            kw.update(y_rot2=self.rot2 if self.cy else 0, rot4=self.rot4 if self.cy else 0)
        return kw

    # ---- index folding ------------------------------------------------------
    def fold(self, lo, hi, axis):
        """Full-cell span -> (lo, hi, crossed, mirrored) on the kept part:
        crossed = the span reached below the cut (its flux doubles back),
        mirrored = the whole span lies below the cut (mapped onto its image)."""
        c = self.cy if axis == "y" else self.cx
        if not c:
            return lo, hi, False, False
        if hi <= c:
            return c - hi, c - lo, False, True
        return max(0, lo - c), max(1, hi - c), lo < c, False

    def foldpt(self, j, axis):
        """Full-cell index -> (kept-part index, mirrored)."""
        c = self.cy if axis == "y" else self.cx
        if not c:
            return j, False
        return (j - c, False) if j >= c else (c - 1 - j, True)

    def rot_guard(self, what, *below):
        """A rotation image of a point is on a different line of the kept
        part, so a region reaching below a rotation cut has no fold."""
        if (self.rot2 or self.rot4) and any(below):
            raise NotImplementedError(f"{what} reaching beyond the rotation cut plane is not supported: place the "
                                      f"region inside the kept half / quadrant")

    def parity(self, comp):
        """(px, py) of a TM component under the cuts (the E vector / H
        pseudo-vector transform times the phase)."""
        px, py = float(self.px), float(self.py)
        if self.rot4:
            if comp not in ("Ez", "Dz", "Dielectric", None):
                raise NotImplementedError(f"get {comp} under Rotate4(Z): a quarter turn maps Hx onto Hy, so the "
                                          f"component has no scalar parity to de-rotate with")
            return 1.0, 1.0
        if self.rot2:
            return {"Ez": (1.0, py), "Dz": (1.0, py), "Hx": (1.0, -py), "Hy": (1.0, -py)}.get(comp, (1.0, 1.0))
        return {"Ez": (px, py), "Dz": (px, py), "Hx": (px, -py), "Hy": (-px, py)}.get(comp, (1.0, 1.0))

    # ---- full-cell reconstruction -------------------------------------------
    def demirror(self, a, parity=(1.0, 1.0)):
        """Kept-part array -> full cell: the missing part is the kept one
        flipped (mirror), reversed in both axes (Rotate2) or turned (Rotate4),
        times its phase."""
        if self.rot4:
            p, q = float(self.py), a
            hi_x = np.concatenate([p * q[::-1, :].T, q], axis=1)
            lo_x = np.concatenate([q[::-1, ::-1], p * q[:, ::-1].T], axis=1)
            return np.concatenate([lo_x, hi_x], axis=0)
        if self.cy:
            lower = a[:, ::-1] if not self.rot2 else a[::-1, ::-1]
            a = np.concatenate([parity[1] * lower, a], axis=1)
        if self.cx:
            a = np.concatenate([parity[0] * a[::-1, :], a], axis=0)
        return a

    def demirror_te_hz(self, a):
        """Kept-part TE Hz -> full cell (nx-1, ny-1): the kept part fills rows
        / columns cx..nx-2 / cy..ny-2, its image (Hz carries -phase) fills
        1..cx-1 / 1..cy-1, and row / column 0 at the outer edge is zero."""
        if self.cy:
            a = np.concatenate([np.zeros((a.shape[0], 1)), -self.py * a[:, ::-1], a], axis=1)
        if self.cx:
            a = np.concatenate([np.zeros((1, a.shape[1])), -self.px * a[::-1, :], a], axis=0)
        return a
