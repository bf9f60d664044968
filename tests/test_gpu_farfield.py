"""GPU far field (emg_farfield) against the oracle (oracle/farfield.py).

Bar: |Eout - Eout_oracle| <= 1e-10 max|Eout_oracle| (FP64; the sums run in
another order), Hout likewise; the dipole's far field also within 1e-3 of
its closed form, and a plane wave through the closed surface radiates
< 3e-3 of its E-term alone.
"""
import numpy as np
import pytest

from emerge_amd._lib import Context, EmgError, ptr
from emerge_amd.farfield import FarFieldSurface, stratton_chu
from oracle import farfield as ff
from tests.test_farfield import K, cube_surface, dipole_fields

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _surface(n):
    m, bt, pos, wn = cube_surface(n)
    s = FarFieldSurface(m, bt)
    return s


def _compare(ctx, s, E, H, T, P, k0):
    Eo, Ho, wn = stratton_chu(E, H, s, T, P, k0, ctx=ctx)
    Er, Hr = ff.stratton_chu_ff(E, H, s.edge_centers, s.weighted_normals, T, P, k0)
    scale = np.abs(Er).max()
    assert np.abs(Eo - Er).max() <= 1e-10 * scale, np.abs(Eo - Er).max() / scale
    assert np.abs(Ho - Hr).max() <= 1e-10 * scale / ff.Z0
    return Eo, Ho


def test_dipole_matches_oracle_and_closed_form(ctx):
    s = _surface(6)
    E, H = dipole_fields(s.edge_centers)
    T = np.linspace(0.2, np.pi - 0.2, 9)
    P = np.full(9, 0.7)
    Eo, _ = _compare(ctx, s, E, H, T, P, K)
    tho = np.stack([np.cos(T) * np.cos(P), np.cos(T) * np.sin(P), -np.sin(T)])
    ref = 1j * ff.Z0 * K * np.sin(T) / (4 * np.pi)
    assert np.abs((Eo * tho).sum(axis=0) - ref).max() <= 1e-3 * np.abs(ref).max()


def test_plane_wave_extinction(ctx):
    s = _surface(6)
    z = s.edge_centers[2]
    E = np.stack([np.exp(-1j * K * z), 0 * z, 0 * z])
    H = np.stack([0 * z, np.exp(-1j * K * z) / ff.Z0, 0 * z])
    T = np.linspace(0.1, np.pi - 0.1, 12)
    P = np.linspace(0, 2 * np.pi, 12)
    Eo, Ho, _ = stratton_chu(E, H, s, T, P, K, ctx=ctx)
    Ee, _, _ = stratton_chu(E, 0 * H, s, T, P, K, ctx=ctx)
    assert np.abs(Eo).max() < 3e-3 * np.abs(Ee).max()


@pytest.mark.parametrize("n,ndir", [(10, 700), (3, 1)])
def test_many_tiles_and_chunks_with_filtered_samples(ctx, n, ndir):
    """Several direction tiles and sample chunks (n = 10: 1,800 samples,
    700 directions), one direction; some samples under the 1e-3 level."""
    s = _surface(n)
    E, H = dipole_fields(s.edge_centers)
    mag = np.sqrt((np.abs(E) ** 2).sum(axis=0))
    low = np.arange(0, E.shape[1], 7)
    E[:, low] *= (0.5e-3 * mag.max() / mag[low])[None, :]
    rng = np.random.default_rng(3)
    T = rng.uniform(0, np.pi, ndir)
    P = rng.uniform(0, 2 * np.pi, ndir)
    print(f"samples={E.shape[1]} directions={ndir}", flush=True)
    _compare(ctx, s, E, H, T, P, 1.3 * K)


def test_rejects_empty_input(ctx):
    z = np.zeros(6)
    with pytest.raises(EmgError):
        ctx.call("emg_farfield", 0, ptr(z), ptr(z), ptr(z), ptr(z), 1, ptr(z), ptr(z), 1.0, ptr(z), ptr(z))


def test_aperture_far_field_from_a_solved_sweep():
    """The whole chain on a solution: WR-90 fed on port 1, ended by a
    first-order absorbing boundary; E and H sampled on the aperture by the
    device (FrequencySweep.interpolate) and summed by emg_farfield, against
    the oracle's interpolation (owner rule included) and Stratton-Chu."""
    from emerge_amd.sweep import C0, EPS0, FrequencySweep
    from oracle import fem
    from tests import wg_case
    c = wg_case.abc_case(1)
    m, f, p = c["mesh"], c["basis"], c["problem"]
    sw = FrequencySweep(m, f, p.eps_r, p.inv_mu_r, c["prod_bcs"])
    freq = 10e9
    k0 = 2 * np.pi * freq / C0
    res = sw.solve([freq], precision="fp64", tol=1e-10)
    x = res.fields[(0, 1)]
    surf = FarFieldSurface(m, c["abc_tris"])
    assert np.all(surf.weighted_normals[2] > 0)      # aperture normals point out of the guide (+z)
    E, H = sw.interpolate(x, *surf.exyz, freq=freq)
    T = np.linspace(0.05, np.pi / 2, 40)
    P = np.full(40, 0.3)
    Eo, Ho, _ = stratton_chu(E, H, surf, T, P, k0)
    allt = np.arange(m.n_tets)
    pts = surf.edge_centers
    Er = np.array(fem.ned2_tet_interp(pts, x, m, f.tet_to_field, allt))
    mu0 = 1 / (C0 * C0 * EPS0)
    Hr = fem.ned2_tet_interp_curl(pts, x, m, f.tet_to_field, allt, 1 / (-1j * 2 * np.pi * freq * mu0 / p.inv_mu_r))
    Efr, Hfr = ff.stratton_chu_ff(Er, Hr, pts, surf.weighted_normals, T, P, k0)
    scale = np.abs(Efr).max()
    print(f"aperture samples={pts.shape[1]} |Eff|max={scale:.3e}", flush=True)
    assert scale > 0
    assert np.abs(Eo - Efr).max() <= 1e-10 * scale, np.abs(Eo - Efr).max() / scale
    assert np.abs(Ho - Hfr).max() <= 1e-10 * scale / ff.Z0


def test_symmetry_plane_reproduces_the_closed_surface(ctx):
    """A z-dipole's x = 0 plane is a magnetic wall ('Hx': E odd in x, H
    even): the half of the cube with x > 0 plus its mirror image gives the
    dipole's far field as the closed cube does."""
    from emerge_amd.farfield import farfield_symmetric
    m, bt, _, _ = cube_surface(6)
    cx = m.nodes[0, m.tris[:, bt]].mean(axis=0)
    half = FarFieldSurface(m, bt[cx > 0], origin=(0.0, 0.0, 0.0))
    E, H = dipole_fields(half.edge_centers)
    T = np.linspace(0.2, np.pi - 0.2, 9)
    P = np.full(9, 0.7)
    Eh, _ = farfield_symmetric(E, H, half, T, P, K, syms=["Hx"], ctx=ctx)
    tho = np.stack([np.cos(T) * np.cos(P), np.cos(T) * np.sin(P), -np.sin(T)])
    ref = 1j * ff.Z0 * K * np.sin(T) / (4 * np.pi)
    err = np.abs((Eh * tho).sum(axis=0) - ref).max() / np.abs(ref).max()
    assert err < 2e-3, err
    E1, _ = farfield_symmetric(E, H, half, T, P, K, syms=None, ctx=ctx)
    assert np.abs((E1 * tho).sum(axis=0) - ref).max() > 0.1 * np.abs(ref).max()   # the half alone is not it
    Ew, _ = farfield_symmetric(E, H, half, T, P, K, syms=["Ex"], ctx=ctx)     # the wrong wall
    assert np.abs((Ew * tho).sum(axis=0) - ref).max() > 0.1 * np.abs(ref).max()


def test_aperture_power_is_the_transmitted_power():
    """Ptot of MWField.farfield through the absorbing aperture of the WR-90
    guide (port 1 fed with 1 W) equals the power not reflected, 1 - |S11|^2,
    to the discretisation: the time-averaged Poynting vector E x H* / 2."""
    from emerge_amd.sweep import FrequencySweep
    from tests import wg_case
    c = wg_case.abc_case(1)
    m, f, p = c["mesh"], c["basis"], c["problem"]
    sw = FrequencySweep(m, f, p.eps_r, p.inv_mu_r, c["prod_bcs"])
    freq = 10e9
    res = sw.solve([freq], precision="fp64", tol=1e-10)
    Eff, Hff, Ptot = sw.farfield(res.fields[(0, 1)], freq, c["abc_tris"], np.array([0.1, 0.5]), np.array([0.0, 0.3]))
    s11 = abs(res.S[0, 0, 0])
    print(f"Ptot={Ptot:.4f} 1-|S11|^2={1 - s11 ** 2:.4f}", flush=True)
    assert abs(Ptot - (1 - s11 ** 2)) < 0.05
    assert np.all(np.isfinite(Eff)) and np.all(np.isfinite(Hff))


def test_cut_and_grid_wrappers_agree_with_farfield():
    from emerge_amd.farfield import arc_on_plane
    from emerge_amd.sweep import FrequencySweep
    from tests import wg_case
    c = wg_case.abc_case(1)
    m, f, p = c["mesh"], c["basis"], c["problem"]
    sw = FrequencySweep(m, f, p.eps_r, p.inv_mu_r, c["prod_bcs"])
    freq = 10e9
    x = sw.solve([freq], precision="fp64", tol=1e-10).fields[(0, 1)]
    ang, E2, H2, P2 = sw.farfield_2d(x, freq, c["abc_tris"], (0, 0, 1), (0, 1, 0), (-90, 90), 7)
    t, ph = arc_on_plane((0, 0, 1), (0, 1, 0), (-90, 90), 7)
    E1, H1, P1 = sw.farfield(x, freq, c["abc_tris"], t, ph)
    np.testing.assert_allclose(E2, E1, rtol=0, atol=1e-13 * np.abs(E1).max())
    assert ang.shape == (7,) and P2 == P1
    T, P, E3, H3, P3 = sw.farfield_3d(x, freq, c["abc_tris"], np.linspace(0, 1, 4), np.linspace(0, 2, 5))
    assert E3.shape == (3, 5, 4) and T.shape == (5, 4)
    E4, _, _ = sw.farfield(x, freq, c["abc_tris"], T.ravel(), P.ravel())
    np.testing.assert_allclose(E3.reshape(3, -1), E4, rtol=0, atol=1e-13 * np.abs(E4).max())
