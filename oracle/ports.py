"""Oracle: port / absorbing boundary conditions, restated.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates (``MW/`` = src/emerge/_emerge/physics/microwave/):
* constants E/const.py:2-6;
* CoordinateSystem.in_local_cs / in_global_basis (E/cs.py:459-499);
* PortBC._qmode / Zmode (MW/microwave_bc.py:221-243);
* RectangularWaveguide (MW/microwave_bc.py:755-895);
* LumpedPort (MW/microwave_bc.py:1018-1163);
* AbsorbingBoundary.get_gamma (MW/microwave_bc.py:291-344).
Pinned by the analytic waveguide closed forms in the tests (beta, S21 phase).
"""
from __future__ import annotations

import numpy as np

C0 = 299792458
Z0 = 376.73031366857
PI = 3.14159265358979323846
EPS0 = 8.854187818814e-12
MU0 = 1 / (C0 * C0 * EPS0)


class CS:
    """Orthonormal frame: columns of `basis` are the local x, y, z axes."""

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
        return (B[0, 0]*xg + B[0, 1]*yg + B[0, 2]*zg, B[1, 0]*xg + B[1, 1]*yg + B[1, 2]*zg,
                B[2, 0]*xg + B[2, 1]*yg + B[2, 2]*zg)

    def in_global_basis(self, x, y, z):
        a, b, c = self.xax, self.yax, self.zax
        return (a[0]*x + b[0]*y + c[0]*z, a[1]*x + b[1]*y + c[1]*z, a[2]*x + b[2]*y + c[2]*z)


class RectangularWaveguide:
    modetype_ = "TE"
    v_integration = False

    def __init__(self, tri_ids, port_number, cs: CS, dims, mode=(1, 0), er=1.0, power=1.0,
                 polarization=1.0):
        self.tri_ids = np.asarray(tri_ids)
        self.port_number = port_number
        self.cs = cs
        self.dims = dims
        self.mode = mode
        self.er = er
        self.power = power
        self._polarization = polarization
        self.active = False

    def modetype(self, k0):
        return "TE"

    def Zmode(self, k0):
        return k0 * 299792458 / self.get_beta(k0) * MU0

    def _qmode(self, k0):
        return np.sqrt(self.Zmode(k0) / Z0)

    def portZ0(self, k0):
        return k0 * 299792458 * MU0 / self.get_beta(k0)

    def get_amplitude(self, k0):
        return np.sqrt(self.power * 4 * Z0 / (self.dims[0] * self.dims[1]))

    def get_beta(self, k0):
        w, h = self.dims
        return np.sqrt(self.er * k0**2 - (np.pi*self.mode[0]/w)**2 - (np.pi*self.mode[1]/h)**2)

    def get_gamma(self, k0):
        return 1j * self.get_beta(k0)

    def port_mode_3d(self, xl, yl, k0):
        w, h = self.dims
        m, n = self.mode
        Ev = self._polarization*self.get_amplitude(k0)*np.cos(np.pi*m*xl/w)*np.cos(np.pi*n*yl/h)
        Eh = self._polarization*self.get_amplitude(k0)*np.sin(np.pi*m*xl/w)*np.sin(np.pi*n*yl/h)
        return self._qmode(k0) * np.array([Eh, Ev, 0*Eh])

    def port_mode_3d_global(self, x, y, z, k0):
        xl, yl, _ = self.cs.in_local_cs(x, y, z)
        Ex, Ey, Ez = self.port_mode_3d(xl, yl, k0)
        return np.array(self.cs.in_global_basis(Ex, Ey, Ez))

    def get_Uinc(self, x, y, z, k0):
        return -2 * 1j * self.get_beta(k0) * self.port_mode_3d_global(x, y, z, k0)


class LumpedPort:
    v_integration = True

    def __init__(self, tri_ids, port_number, width, height, direction, power=1.0, Z0_port=50.0):
        self.tri_ids = np.asarray(tri_ids)
        self.port_number = port_number
        self.width = width
        self.height = height
        self.direction = np.asarray(direction, float) / np.linalg.norm(direction)
        self.power = power
        self.Z0 = Z0_port
        self.active = False
        self.vintline = []

    def modetype(self, k0):
        return "TEM"

    @property
    def surfZ(self):
        return self.Z0 * self.width / self.height

    @property
    def voltage(self):
        return np.sqrt(2 * self.power * self.Z0)

    def get_beta(self, k0):
        return k0

    def get_gamma(self, k0):
        return 1j * k0 * Z0 / self.surfZ

    def port_mode_3d_global(self, x, y, z, k0):
        on = np.ones_like(x)
        return np.array([self.direction[0]*on, self.direction[1]*on, self.direction[2]*on])

    def get_Uinc(self, x, y, z, k0):
        Emag = -1j * 2 * k0 * self.voltage / self.height * (Z0 / self.surfZ)
        return Emag * self.port_mode_3d_global(x, y, z, k0)


class AbsorbingBoundary:
    o2coeffs = {"A": (1.0, -0.5), "B": (1.00023, -0.51555), "C": (1.03084, -0.73631),
                "D": (1.06103, -0.84883), "E": (1.12500, -1.00000)}

    def __init__(self, tri_ids, order=1, abctype="B"):
        self.tri_ids = np.asarray(tri_ids)
        self.order = order
        self.abctype = abctype

    def get_gamma(self, k0):
        if self.order == 1:
            return 1j * k0
        return 1j * k0 * self.o2coeffs[self.abctype][0]
