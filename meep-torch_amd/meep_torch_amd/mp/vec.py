"""meep-style vectors and constants (meep_torch/mp/vec.py, meep_torch/mp/__init__.py)."""
import math

inf = float("inf")

# field components (meep names; TM = Ez/Hx/Hy, TE = Hz/Ex/Ey)
Ex, Ey, Ez, Hx, Hy, Hz, Dz = "Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Dz"
Dielectric = "Dielectric"
TM_COMPONENTS = ("Ez", "Hx", "Hy", "Dz")
TE_COMPONENTS = ("Hz", "Ex", "Ey")

# directions and sides for boundary layers
X, Y, Z = 0, 1, 2
ALL = -1
Low, High = 0, 1


class Vector3:
    __slots__ = ("x", "y", "z")

    def __init__(self, x=0.0, y=0.0, z=0.0):
        self.x, self.y, self.z = float(x), float(y), float(z)

    def __iter__(self):
        return iter((self.x, self.y, self.z))

    def __add__(self, o):
        return Vector3(self.x + o.x, self.y + o.y, self.z + o.z)

    def __sub__(self, o):
        return Vector3(self.x - o.x, self.y - o.y, self.z - o.z)

    def __neg__(self):
        return Vector3(-self.x, -self.y, -self.z)

    def __eq__(self, o):
        return isinstance(o, Vector3) and (self.x, self.y, self.z) == (o.x, o.y, o.z)

    def __hash__(self):
        return hash((self.x, self.y, self.z))

    def scale(self, s):
        return Vector3(self.x * s, self.y * s, self.z * s)

    def norm(self):
        return math.sqrt(self.x ** 2 + self.y ** 2 + self.z ** 2)

    def unit(self):
        n = self.norm()
        return Vector3(self.x / n, self.y / n, self.z / n)

    def __repr__(self):
        return f"Vector3({self.x}, {self.y}, {self.z})"


def pol_for(component):
    """'tm' for Ez-family sources, 'te' for Hz-family (meep_torch/mp/sources.py:43-48)."""
    if component in ("Ez", "Hx", "Hy", "Dz"):
        return "tm"
    if component in ("Hz", "Ex", "Ey"):
        return "te"
    raise ValueError(f"unsupported component {component!r}")
