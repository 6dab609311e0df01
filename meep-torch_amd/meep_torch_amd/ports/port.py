"""Port planes of the S-parameter drivers (meep_torch/ports/port.py:1-42)."""
from dataclasses import dataclass

# cells between a port's monitor plane and its TFSF source sheets (TM); the
# TE source sits TE_SRC_OFFSET cells upstream (its weakly guided mode has a
# long evanescent tail) (ports/port.py:8, ports/mode_source.py:19)
SRC_OFFSET = 3
TE_SRC_OFFSET = 12


@dataclass(frozen=True)
class Port:
    """A monitor line normal to ``axis`` at ``index``, spanning [lo, hi) of
    the transverse axis; ``direction`` is the launch direction (+1 / -1)."""
    axis: str
    index: int
    lo: int
    hi: int
    direction: int
    n_modes: int = 1


def validate_ports(ports, grid, npml: int) -> None:
    """ports/port.py:24-42: axis, direction, mode count, span inside the grid
    and off the PML, plane at least SRC_OFFSET cells inside the PML."""
    for q, p in enumerate(ports):
        if p.axis not in ("x", "y"):
            raise ValueError(f"port {q}: axis must be 'x' or 'y', got {p.axis!r}")
        if p.direction not in (1, -1):
            raise ValueError(f"port {q}: direction must be +1 or -1, got {p.direction}")
        if p.n_modes < 1:
            raise ValueError(f"port {q}: n_modes must be >= 1, got {p.n_modes}")
        n_along, n_across = (grid.nx, grid.ny) if p.axis == "x" else (grid.ny, grid.nx)
        if not p.lo < p.hi:
            raise ValueError(f"port {q}: empty span lo={p.lo}, hi={p.hi}")
        if p.lo < 0 or p.hi > n_across:
            raise ValueError(f"port {q}: span [{p.lo},{p.hi}) outside grid (n={n_across})")
        if not npml + SRC_OFFSET <= p.index < n_along - npml - SRC_OFFSET:
            raise ValueError(f"port {q}: plane index={p.index} inside/adjacent to PML "
                             f"(need {npml + SRC_OFFSET} <= index < {n_along - npml - SRC_OFFSET})")
        if p.lo < npml or p.hi > n_across - npml:
            raise ValueError(f"port {q}: span [{p.lo},{p.hi}) reaches into PML (npml={npml})")
