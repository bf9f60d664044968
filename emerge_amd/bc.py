"""Boundary conditions of the frequency-domain path (host-side scalars).

Mirror the reference classes (``MW/`` = src/emerge/_emerge/physics/microwave/):
PEC (MW/microwave_bc.py:116-129), RectangularWaveguide (:755-895),
LumpedPort (:1018-1163) and AbsorbingBoundary (:282-344), with the port base
behaviour of PortBC (:175-280).  Faces are given as triangle index arrays of
the mesh (gmsh face tags are not available).  Each Robin BC supplies
``get_gamma(k0)`` and, when excited, an incident field that factors as
``get_Uinc(x, y, z, k0) = uinc_factor(k0) * uinc_shape(x, y, z)`` so the
device evaluates the forcing integrals once per port.
"""
from __future__ import annotations

import numpy as np

C0 = 299792458
Z0 = 376.73031366857
EPS0 = 8.854187818814e-12
MU0 = 1 / (C0 * C0 * EPS0)


class CoordinateSystem:
    """Orthonormal frame; ``in_local_cs`` / ``in_global_basis`` as E/cs.py:459-499."""

    def __init__(self, xax, yax, zax, origin):
        self.xax = np.asarray(xax, float) / np.linalg.norm(xax)
        self.yax = np.asarray(yax, float) / np.linalg.norm(yax)
        self.zax = np.asarray(zax, float) / np.linalg.norm(zax)
        self.origin = np.asarray(origin, float)
        self._basis = np.array([self.xax, self.yax, self.zax]).T
        self._basis_inv = np.linalg.pinv(self._basis)

    def in_local_cs(self, x, y, z):
        B = self._basis_inv
        xg, yg, zg = x - self.origin[0], y - self.origin[1], z - self.origin[2]
        return (B[0, 0] * xg + B[0, 1] * yg + B[0, 2] * zg,
                B[1, 0] * xg + B[1, 1] * yg + B[1, 2] * zg,
                B[2, 0] * xg + B[2, 1] * yg + B[2, 2] * zg)

    def in_global_basis(self, x, y, z):
        a, b, c = self.xax, self.yax, self.zax
        return (a[0] * x + b[0] * y + c[0] * z, a[1] * x + b[1] * y + c[1] * z,
                a[2] * x + b[2] * y + c[2] * z)


class BoundaryCondition:
    def __init__(self, tri_ids):
        self.tri_ids = np.asarray(tri_ids, dtype=np.int64)


class PEC(BoundaryCondition):
    """Perfect conductor: its edge and face DOFs are eliminated."""


class RobinBC(BoundaryCondition):
    excited = False
    is_port = False

    def get_gamma(self, k0: float) -> complex:
        raise NotImplementedError


class AbsorbingBoundary(RobinBC):
    """First- or second-order absorbing boundary (MW/microwave_bc.py:291-344;
    order 2 with the coefficient sets A-E, default order 2 type B as the
    reference).  Order 2 adds the surface curl/div correction with factor
    1j c2 / k0 (ASM/assembler.py:342-348)."""
    o2coeffs = {"A": (1.0, -0.5), "B": (1.00023, -0.51555), "C": (1.03084, -0.73631),
                "D": (1.06103, -0.84883), "E": (1.12500, -1.00000)}

    def __init__(self, tri_ids, order: int = 2, abctype: str = "B"):
        super().__init__(tri_ids)
        if order not in (1, 2):
            raise ValueError("order must be 1 or 2")
        if abctype not in self.o2coeffs:
            raise ValueError(f"abctype must be one of {sorted(self.o2coeffs)}")
        self.order = order
        self.abctype = abctype

    def get_gamma(self, k0):
        if self.order == 1:
            return 1j * k0
        return 1j * k0 * self.o2coeffs[self.abctype][0]

    def abc2_coefficient(self, k0):
        """Factor of the second-order correction term."""
        return 1j * self.o2coeffs[self.abctype][1] / k0


class PortBC(RobinBC):
    excited = True
    is_port = True

    def __init__(self, tri_ids, port_number: int, power: float = 1.0):
        super().__init__(tri_ids)
        self.port_number = port_number
        self.power = power
        self.active = False

    def modetype(self, k0):
        return "TEM"

    def get_beta(self, k0):
        return k0

    def Zmode(self, k0):
        mt = self.modetype(k0)
        if mt == "TEM":
            return Z0
        if mt == "TE":
            return k0 * 299792458 / self.get_beta(k0) * MU0
        return self.get_beta(k0) / (k0 * 299792458 * EPS0)

    def _qmode(self, k0):
        return np.sqrt(self.Zmode(k0) / Z0)

    def get_Uinc(self, x, y, z, k0):
        return self.uinc_factor(k0) * self.uinc_shape(x, y, z)


class RectangularWaveguide(PortBC):
    """TE_mn port; local frame origin at the face centre, x along the long side."""

    def __init__(self, tri_ids, port_number: int, cs: CoordinateSystem, dims, mode=(1, 0),
                 er: float = 1.0, power: float = 1.0):
        super().__init__(tri_ids, port_number, power)
        self.cs = cs
        self.dims = tuple(dims)
        self.mode = tuple(mode)
        self.er = er
        self._polarization = 1.0

    def modetype(self, k0):
        return "TE"

    def portZ0(self, k0):
        return k0 * 299792458 * MU0 / self.get_beta(k0)

    def get_amplitude(self, k0=None):
        return np.sqrt(self.power * 4 * Z0 / (self.dims[0] * self.dims[1]))

    def get_beta(self, k0):
        w, h = self.dims
        return np.sqrt(self.er * k0 ** 2 - (np.pi * self.mode[0] / w) ** 2 - (np.pi * self.mode[1] / h) ** 2)

    def get_gamma(self, k0):
        return 1j * self.get_beta(k0)

    def mode_shape(self, x, y, z):
        """Mode field without the _qmode(k0) factor, global components (3, n)."""
        xl, yl, _ = self.cs.in_local_cs(x, y, z)
        w, h = self.dims
        m, n = self.mode
        amp = self._polarization * self.get_amplitude()
        ev = amp * np.cos(np.pi * m * xl / w) * np.cos(np.pi * n * yl / h)
        eh = amp * np.sin(np.pi * m * xl / w) * np.sin(np.pi * n * yl / h)
        return np.array(self.cs.in_global_basis(eh, ev, 0 * eh))

    def mode_factor(self, k0):
        return self._qmode(k0)

    def port_mode_3d_global(self, x, y, z, k0):
        return self.mode_factor(k0) * self.mode_shape(x, y, z)

    def uinc_shape(self, x, y, z):
        return self.mode_shape(x, y, z)

    def uinc_factor(self, k0):
        return -2 * 1j * self.get_beta(k0) * self._qmode(k0)


class LumpedPort(PortBC):
    """Uniform-field sheet port of impedance Z0_port (MW/microwave_bc.py:1018-1163);
    its S-parameters come from the voltage along lines over the port height."""

    v_integration = True

    def __init__(self, tri_ids, port_number: int, width: float, height: float, direction,
                 power: float = 1.0, Z0_port: float = 50.0):
        super().__init__(tri_ids, port_number, power)
        self.width = width
        self.height = height
        self.direction = np.asarray(direction, float) / np.linalg.norm(direction)
        self.Z0 = Z0_port

    @property
    def surfZ(self):
        return self.Z0 * self.width / self.height

    @property
    def voltage(self):
        return np.sqrt(2 * self.power * self.Z0)

    def get_gamma(self, k0):
        return 1j * k0 * Z0 / self.surfZ

    def mode_shape(self, x, y, z):
        on = np.ones_like(np.asarray(x, dtype=float))
        return np.array([self.direction[0] * on, self.direction[1] * on, self.direction[2] * on])

    def mode_factor(self, k0):
        return 1.0

    def port_mode_3d_global(self, x, y, z, k0):
        return self.mode_shape(x, y, z)

    def uinc_shape(self, x, y, z):
        return self.mode_shape(x, y, z)

    def uinc_factor(self, k0):
        return -1j * 2 * k0 * self.voltage / self.height * (Z0 / self.surfZ)
