"""Pins of the boundary-mode oracle (oracle/modal.py) by closed forms.

The port face of a WR-90 guide must carry the TE10 mode with the analytic
propagation constant (converging as the face is refined) and, once
normalised to unit power, the analytic TE10 field; a parallel-plate face
must carry the TEM mode with beta = k0 sqrt(eps_r); the eigenvalues must not
change under a rigid motion of the face.
"""
from __future__ import annotations

import numpy as np
import pytest

from emerge_amd.mesh import Mesh3D, kuhn_box
from oracle import modal as OM
from oracle import ports as oports

A_WR, B_WR = 22.86e-3, 10.16e-3
F = 10e9
K0 = 2 * np.pi * F / OM.C0


def _face(nx, ny, size, pec_planes, rot=None, shift=None):
    nodes, tets = kuhn_box(nx, ny, 2, size=(size[0], size[1], 4e-3))
    if rot is not None:
        nodes = rot @ nodes + shift[:, None]
    m = Mesh3D(nodes, tets)
    z = np.array([0.0, 0.0, 1.0]) if rot is None else rot[:, 2]
    # port face: z_local = 0; conductor faces: the given planes of the box
    loc = nodes if rot is None else rot.T @ (nodes - shift[:, None])
    tri_loc = loc[:, m.tris]
    on = lambda ax, val: np.all(np.abs(tri_loc[ax] - val) < 1e-12, axis=0)
    port = np.nonzero(on(2, 0.0))[0]
    pec = np.concatenate([np.nonzero(on(ax, val))[0] for ax, val in pec_planes])
    o = np.zeros(3) if shift is None else shift
    xax = np.array([1.0, 0, 0]) if rot is None else rot[:, 0]
    yax = np.array([0, 1.0, 0]) if rot is None else rot[:, 1]
    cs = oports.CS(xax, yax, z, o)
    n = port.size
    eye = np.eye(3)[:, :, None] * np.ones(n)
    return m, port, pec, cs, eye.astype(complex)


def _wr90(nx, ny, rot=None, shift=None):
    planes = [(0, 0.0), (0, A_WR), (1, 0.0), (1, B_WR)]
    return _face(nx, ny, (A_WR, B_WR), planes, rot, shift)


def _beta_te10():
    return np.sqrt(K0 ** 2 - (np.pi / A_WR) ** 2)


@pytest.fixture(scope="module")
def wr90_coarse():
    m, port, pec, cs, eye = _wr90(11, 5)
    modes, surf = OM.modal_analysis(m.nodes, m.tris, port, cs, eye, eye, pec, K0)
    return m, port, cs, modes, surf


def test_te10_beta_converges(wr90_coarse):
    *_, modes, _ = wr90_coarse
    assert len(modes) == 1 and modes[0]["modetype"] == "TE"   # TE20 and TE01 are cut off at 10 GHz
    e1 = abs(modes[0]["beta"] / _beta_te10() - 1)
    m, port, pec, cs, eye = _wr90(20, 9)
    fine, _ = OM.modal_analysis(m.nodes, m.tris, port, cs, eye, eye, pec, K0, nmodes=2, direct=False)
    e2 = min(abs(md["beta"] / _beta_te10() - 1) for md in fine)
    assert e1 < 2e-3 and e2 < e1 / 3, (e1, e2)


def test_te10_normalised_field_matches_closed_form(wr90_coarse):
    """Unit-power normalisation (power flux of E and the curl-based H) and the
    sin(pi x / a) shape: the normalised modal field against the analytic TE10
    port field (amplitude sqrt(4 P Z_TE / (a b))) at interior points."""
    m, port, cs, modes, surf = wr90_coarse
    mp = OM.ModalPortOracle(port, 1, cs, modes, surf, F)
    rw = oports.RectangularWaveguide(port, 1, oports.CS([1, 0, 0], [0, 1, 0], [0, 0, 1],
                                                       [A_WR / 2, B_WR / 2, 0.0]), (A_WR, B_WR))
    g = np.random.default_rng(0)
    x = np.concatenate([[A_WR / 2], g.uniform(0.1, 0.9, 200) * A_WR])
    y = np.concatenate([[B_WR / 2], g.uniform(0.1, 0.9, 200) * B_WR])
    z = np.zeros_like(x)
    Em = mp.port_mode_3d_global(x, y, z, K0)
    Ea = rw.port_mode_3d_global(x, y, z, K0)
    s = np.sign(np.real(np.sum(Em[1] * Ea[1])))
    centre = abs(s * Em[1, 0] / Ea[1, 0] - 1)           # amplitude of the unit-power mode
    rms = np.sqrt(np.mean(np.abs(s * Em - Ea) ** 2)) / np.max(np.abs(Ea))
    assert centre < 3e-3 and rms < 1e-2, (centre, rms)


def test_parallel_plate_tem():
    """Conductors on y = 0 and y = h only: the TEM mode (uniform E_y) is in
    the discrete space, so beta = k0 sqrt(eps_r) to rounding."""
    er = 2.2
    m, port, pec, cs, eye = _face(6, 4, (8e-3, 2e-3), [(1, 0.0), (1, 2e-3)])
    modes, surf = OM.modal_analysis(m.nodes, m.tris, port, cs, er * eye, eye, pec, K0, tem=True)
    betas = np.array([md["beta"] for md in modes])
    best = betas[np.argmin(np.abs(betas - K0 * np.sqrt(er)))]
    assert abs(best / (K0 * np.sqrt(er)) - 1) < 1e-9, betas


def test_rigid_motion_invariance(wr90_coarse):
    *_, modes, _ = wr90_coarse
    th = 0.7
    rot = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    rot = rot @ np.array([[np.cos(0.3), -np.sin(0.3), 0], [np.sin(0.3), np.cos(0.3), 0], [0, 0, 1]])
    m, port, pec, cs, eye = _wr90(11, 5, rot=rot, shift=np.array([0.3, -1.2, 2.0]))
    moved, _ = OM.modal_analysis(m.nodes, m.tris, port, cs, eye, eye, pec, K0)
    assert abs(moved[0]["beta"] / modes[0]["beta"] - 1) < 1e-9


def _slab_case():
    """Parallel plates (conductors y = 0, y = h, open sides) loaded with a
    dielectric slab eps1 on x < a and eps2 beyond: the E_y(x) mode's beta
    solves kx1 tan(kx1 a) = ka tanh(ka (W - a)), kx1^2 = k0^2 eps1 - beta^2,
    ka^2 = beta^2 - k0^2 eps2 (magnetic side walls)."""
    from scipy.optimize import brentq
    W, hh, a, e1, e2 = 8e-3, 2e-3, 3e-3, 4.0, 1.0
    k0 = 2 * np.pi * 10e9 / OM.C0

    def char(beta):
        kx1 = np.sqrt(k0 ** 2 * e1 - beta ** 2)
        ka = np.sqrt(beta ** 2 - k0 ** 2 * e2)
        return kx1 * np.tan(kx1 * a) - ka * np.tanh(ka * (W - a))

    bs = np.linspace(k0 * np.sqrt(e2) * 1.0001, k0 * np.sqrt(e1) * 0.9999, 2000)
    v = char(bs)
    roots = [brentq(char, bs[i], bs[i + 1]) for i in range(len(bs) - 1)
             if np.sign(v[i]) != np.sign(v[i + 1]) and abs(v[i]) < 1e3 and abs(v[i + 1]) < 1e3]
    nodes, tets = kuhn_box(16, 2, 1, size=(W, hh, 1e-3))
    m = Mesh3D(nodes, tets)
    tv = m.nodes[:, m.tris]
    on = lambda ax, val: np.all(np.abs(tv[ax] - val) < 1e-12, axis=0)
    port = np.nonzero(on(2, 0.0))[0]
    pec = np.nonzero(on(1, 0.0) | on(1, hh))[0]
    eps_t = np.where(m.nodes[:, m.tets].mean(axis=1)[0] < a, e1, e2)
    cs = ([1, 0, 0], [0, 1, 0], [0, 0, 1], [W / 2, hh / 2, 0])
    return dict(mesh=m, port=port, pec=pec, eps_t=eps_t, cs=cs, k0=k0, exact=np.array(roots))


def test_slab_loaded_parallel_plate():
    """Inhomogeneous filling: the face's fundamental mode against the root
    of the slab-loaded parallel-plate characteristic equation."""
    c = _slab_case()
    assert c["exact"].size == 1
    m, port = c["mesh"], c["port"]
    er = (np.eye(3)[:, :, None] * c["eps_t"][m.tri_to_tet[0, port]][None, None, :]).astype(complex)
    ur = (np.eye(3)[:, :, None] * np.ones(port.size)).astype(complex)
    modes, _ = OM.modal_analysis(m.nodes, m.tris, port, oports.CS(*c["cs"]), er, ur, c["pec"], c["k0"])
    beta = max(np.real(md["beta"]) for md in modes)
    assert abs(beta / c["exact"][0] - 1) < 1e-3, (beta / c["k0"], c["exact"] / c["k0"])
