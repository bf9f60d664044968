"""GPU surface terms and Krylov solves against the oracle (SuperLU, reference options).

Bar: surface/forcing values within 1e-12 relative; every solve reaches the
FP64 relative residual ||b - A x|| / ||b|| <= 1e-8 (recomputed here with the
oracle's matrix); S-parameters within |dS| <= 1e-6 of the SuperLU solution.
"""
import numpy as np
import pytest

from emerge_amd.solver import SolverHIP
from emerge_amd.sweep import FrequencySweep
from oracle import robin
from tests import wg_case

pytestmark = pytest.mark.gpu

F0 = 10e9


@pytest.fixture(scope="module")
def sweep():
    c = wg_case.case()
    p = c["problem"]
    return FrequencySweep(c["mesh"], c["basis"], p.eps_r, p.inv_mu_r, c["prod_bcs"])


def test_triangle_terms_match_oracle(sweep):
    c = wg_case.case()
    m = c["mesh"]
    n = sweep.ctx  # noqa
    ntot = sum(b.tri_ids.shape[0] for b in sweep.robin)
    mat = np.empty((ntot, 64))
    force = np.empty((ntot, 8), dtype=np.complex128)
    from emerge_amd._lib import ptr
    sweep.ctx.call("emg_tri_terms_get", ptr(mat), ptr(force))
    pts = robin.gaus_quad_tri(4)
    port = sweep.ports[0]
    off = sweep.tri_off[sweep.robin.index(port)]
    from emerge_amd.sweep import dunavant_points
    P = dunavant_points(m.nodes, m.tris, port.tri_ids)
    for k, t in enumerate(port.tri_ids[:40]):
        shp = port.uinc_shape(P[k, :, 0], P[k, :, 1], P[k, :, 2])
        Bo, bo = robin.ned2_tri_stiff_force(m.nodes[:, m.tris[:, t]], 1.0, shp, pts)
        np.testing.assert_allclose(mat[off + k], Bo.real.ravel(), rtol=0, atol=1e-12 * np.abs(Bo).max())
        np.testing.assert_allclose(force[off + k], bo, rtol=0, atol=1e-12 * np.abs(bo).max())


def test_reduced_operator_matches_oracle(sweep):
    """E_red - k0^2 B_red + sum_bc gamma_bc S_bc built from the device pieces
    equals the oracle's K[ids, ids] (K = E - k0^2 B + B_p)."""
    from scipy.sparse import csr_matrix
    from emerge_amd._lib import ptr
    o = wg_case.oracle_solution(F0)
    # the sweep keeps the oracle's DOFs, in its own (spatial) order
    np.testing.assert_array_equal(np.sort(sweep.solve_ids), o["ids"])
    ids = sweep.solve_ids
    Kr = o["K"].tocsr()[ids][:, ids].tocsr()
    n = ids.shape[0]
    sval = np.empty(sweep.n_sentries)
    sweep.ctx.call("emg_surface_get", ptr(sval))
    E, B = sweep.ctx.values_get()
    full_ip, full_ix = sweep.dm.pattern
    N = sweep.basis.n_field
    Er = csr_matrix((E, full_ix, full_ip), shape=(N, N))[ids][:, ids]
    Br = csr_matrix((B, full_ix, full_ip), shape=(N, N))[ids][:, ids]
    k0 = o["k0"]
    gam = np.array([b.get_gamma(k0) for b in sweep.robin])
    rows = np.repeat(np.arange(n), np.diff(sweep.s_ptr))
    S = csr_matrix((gam[sweep.s_bc] * sval, (rows, sweep.s_col)), shape=(n, n))
    A = (Er - k0 ** 2 * Br + S).tocsr()
    D = (A - Kr).tocsr()
    assert np.abs(D.data).max(initial=0) <= 1e-12 * np.abs(Kr.data).max()
    # the reduced pattern holds every entry of the oracle's reduced matrix
    ip, ix = sweep.reduced_pattern()
    P = csr_matrix((np.ones(ix.shape[0]), ix, ip), shape=(n, n))
    assert (abs(Kr).sign() - abs(Kr).sign().multiply(P)).nnz == 0


def test_forcing_matches_oracle(sweep):
    o = wg_case.oracle_solution(F0)
    fv = sweep.forcing()
    ids = sweep.solve_ids
    for j, p in enumerate(sweep.ports):
        ref = o["pv"][p.port_number][ids] / p.uinc_factor(o["k0"])
        np.testing.assert_allclose(fv[j], ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sweep_single_frequency_matches_superlu(sweep, precision):
    print(f"sweep precision={precision} f={F0}", flush=True)
    o = wg_case.oracle_solution(F0)
    res = sweep.solve([F0], precision=precision, tol=1e-8, fblock=2, restart=100, max_cycles=400, method="gmres",
                      inner_rel32=1e-5)
    ids = o["ids"]
    K = o["K"].tocsr()
    for p in sweep.ports:
        x = res.fields[(0, p.port_number)]
        b = o["b"] + o["pv"][p.port_number]
        rr = np.linalg.norm((b - K @ x)[ids]) / np.linalg.norm(b[ids])
        assert rr <= 1e-8, (p.port_number, rr)
    fields = {p.port_number: res.fields[(0, p.port_number)] for p in sweep.ports}
    S = wg_case.s_from_fields(fields, F0)
    assert np.abs(S - o["S"]).max() <= 1e-6, np.abs(S - o["S"]).max()
    # device-side S extraction (port overlaps) vs the oracle's extraction
    assert np.abs(res.S[0] - o["S"]).max() <= 1e-6, np.abs(res.S[0] - o["S"]).max()
    print(f"iters={res.iters.tolist()} relres={res.relres.tolist()} dS={np.abs(S - o['S']).max():.2e} "
          f"dS_dev={np.abs(res.S[0] - o['S']).max():.2e}")


def test_sweep_batched_mixed_several_frequencies(sweep):
    freqs = np.array([8e9, 9.5e9, 11e9, 12e9])
    res = sweep.solve(freqs, precision="mixed", tol=1e-8, fblock=8, restart=100, max_cycles=400, method="gmres",
                      inner_rel32=1e-5)
    assert np.all(res.relres <= 1e-8), res.relres
    for i, f in enumerate(freqs):
        o = wg_case.oracle_solution(float(f))
        fields = {p.port_number: res.fields[(i, p.port_number)] for p in sweep.ports}
        S = wg_case.s_from_fields(fields, float(f))
        assert np.abs(S - o["S"]).max() <= 1e-6, (f, np.abs(S - o["S"]).max())
        assert np.abs(res.S[i] - o["S"]).max() <= 1e-6, (f, np.abs(res.S[i] - o["S"]).max())


def test_solver_plugin_matches_superlu():
    o = wg_case.oracle_solution(F0)
    ids = o["ids"]
    A = o["K"].tocsc()[np.ix_(ids, ids)]
    b = (o["b"] + o["pv"][1])[ids]
    s = SolverHIP("", precision="fp64", tol=1e-10)
    x, rep = s.solve(A, b, None, False, 0)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-10
    ref = o["fields"][1][ids]
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-6
    x2, _ = s.solve(A, (o["b"] + o["pv"][2])[ids], None, True, 0)
    assert np.linalg.norm(x2 - o["fields"][2][ids]) / np.linalg.norm(o["fields"][2][ids]) <= 1e-6
    sm = SolverHIP("", precision="mixed", tol=1e-10)
    x3, rep3 = sm.solve(A, b, None, False, 0)
    assert np.linalg.norm(b - A @ x3) / np.linalg.norm(b) <= 1e-10
    assert rep3.aux["precision"] == "mixed"


def test_idr_rejects_fp32(sweep):
    from emerge_amd._lib import EmgError
    with pytest.raises(EmgError):
        sweep.solve([10e9], precision="mixed", method="idr", restart=100, max_cycles=1)


@pytest.mark.parametrize("precision,s", [("fp64", 4), ("fp64", 8)])
def test_sweep_idr_matches_superlu(sweep, precision, s):
    print(f"idr precision={precision} s={s}", flush=True)
    freqs = np.array([8e9, 10e9])
    res = sweep.solve(freqs, precision=precision, tol=1e-8, fblock=4, restart=30000, max_cycles=8,
                      method="idr", idr_s=s, inner_rel32=1e-4)
    assert np.all(res.relres <= 1e-8), res.relres
    for i, f in enumerate(freqs):
        o = wg_case.oracle_solution(float(f))
        assert np.abs(res.S[i] - o["S"]).max() <= 1e-6, (f, np.abs(res.S[i] - o["S"]).max())
    print(f"idr iters={res.iters.tolist()} relres={res.relres.max():.2e}")


@pytest.mark.parametrize("precision,fblock", [("fp64", 1), ("fp64", 4), ("mixed", 4), ("mixed", 16), ("m32", 4)])
def test_sweep_cocg_matches_superlu(sweep, precision, fblock):
    print(f"cocg precision={precision} fblock={fblock}", flush=True)
    freqs = np.array([8e9, 10e9, 12e9]) if fblock > 1 else np.array([10e9])
    res = sweep.solve(freqs, precision=precision, tol=1e-8, fblock=fblock, restart=20000, max_cycles=30,
                      method="cocg", inner_rel32=1e-2)
    assert np.all(res.relres <= 1e-8), res.relres
    for i, f in enumerate(freqs):
        o = wg_case.oracle_solution(float(f))
        assert np.abs(res.S[i] - o["S"]).max() <= 1e-6, (f, np.abs(res.S[i] - o["S"]).max())
    print(f"cocg iters={res.iters.tolist()} relres={res.relres.max():.2e}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solver_plugin_cocg(precision):
    """SolverHIP (generic CSR) with COCG on the symmetric part of A."""
    o = wg_case.oracle_solution(F0)
    ids = o["ids"]
    A = o["K"].tocsc()[np.ix_(ids, ids)]
    b = (o["b"] + o["pv"][1])[ids]
    s = SolverHIP("", precision=precision, tol=1e-10, method="cocg", restart=20000, max_cycles=40)
    x, rep = s.solve(A, b, None, False, 0)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-10
    ref = o["fields"][1][ids]
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-6
    assert rep.aux["method"] == "cocg"


@pytest.mark.parametrize("precision", ["fp64", "mixed", "m32"])
def test_cavity_lumped_port_matches_superlu(precision):
    """Lumped port (voltage-line S11) in a lossy PEC cavity: residual and S11
    against SuperLU with the oracle's S extraction."""
    from tests import cavity_case
    c = cavity_case.case()
    sw = FrequencySweep(c["mesh"], c["basis"], c["eps"], c["inv_mu"], c["prod_bcs"])
    freqs = np.array([6e9, 9e9])
    res = sw.solve(freqs, precision=precision, tol=1e-8, fblock=2, restart=20000, max_cycles=40,
                   method="cocg", inner_rel32=1e-2)
    assert np.all(res.relres <= 1e-8), res.relres
    for i, f in enumerate(freqs):
        o = cavity_case.oracle_solution(float(f))
        assert np.abs(res.S[i] - o["S"]).max() <= 1e-6, (f, res.S[i], o["S"])
        x = res.fields[(i, 1)]
        ids = o["ids"]
        ref = o["fields"][1][ids]
        assert np.linalg.norm(x[ids] - ref) / np.linalg.norm(ref) <= 1e-6
    print(f"cavity S11={res.S[:, 0, 0]} iters={res.iters.ravel().tolist()}")


def test_abc2_triangle_terms_match_oracle():
    """Device second-order ABC correction per triangle (without cf) vs the
    oracle restatement of _abc_order_2_terms."""
    from emerge_amd._lib import ptr
    from oracle import abc2
    c = wg_case.abc_case(2)
    sw = FrequencySweep(c["mesh"], c["basis"], c["problem"].eps_r, c["problem"].inv_mu_r, c["prod_bcs"])
    kinds = [k for k, _ in sw.surface]
    ia = kinds.index("abc2")
    n_all = sum(b.tri_ids.shape[0] for _, b in sw.surface)
    mat = np.empty(64 * n_all)
    sw.ctx.call("emg_tri_terms_get", ptr(mat), None)
    m = c["mesh"]
    t = c["abc_tris"]
    got = mat[64 * sw.tri_off[ia]: 64 * (sw.tri_off[ia] + t.shape[0])].reshape(-1, 8, 8)
    for k, it in enumerate(t):
        ref = abc2.abc_order_2_terms(m.nodes[:, m.tris[:, it]], 1.0).real
        assert np.abs(got[k] - ref).max() <= 1e-12 * np.abs(ref).max(), it


@pytest.mark.parametrize("order", [1, 2])
def test_waveguide_abc_termination_matches_superlu(order):
    """WR-90 guide closed by an absorbing boundary (order 1 or 2, type B):
    residual, field and S11 against SuperLU and the oracle's extraction."""
    c = wg_case.abc_case(order)
    sw = FrequencySweep(c["mesh"], c["basis"], c["problem"].eps_r, c["problem"].inv_mu_r, c["prod_bcs"])
    freqs = np.array([9e9, 11e9])
    res = sw.solve(freqs, precision="mixed", tol=1e-8, fblock=2, restart=20000, max_cycles=40,
                   method="cocg", inner_rel32=1e-2)
    assert np.all(res.relres <= 1e-8), res.relres
    for i, f in enumerate(freqs):
        o = wg_case.abc_oracle_solution(float(f), order)
        assert np.abs(res.S[i] - o["S"]).max() <= 1e-6, (f, res.S[i], o["S"])
        ids = o["ids"]
        ref = o["fields"][1][ids]
        assert np.linalg.norm(res.fields[(i, 1)][ids] - ref) / np.linalg.norm(ref) <= 1e-6
    print(f"abc order {order}: S11={res.S[:, 0, 0]} iters={res.iters.ravel().tolist()}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sweep_cocg_jacobi_only_matches_superlu(sweep, precision):
    """COCG with Jacobi alone (no auxiliary gradient space)."""
    print(f"cocg jacobi precision={precision}", flush=True)
    freqs = np.array([9e9, 11e9])
    res = sweep.solve(freqs, precision=precision, tol=1e-8, fblock=2, restart=20000, max_cycles=30,
                      method="cocg", inner_rel32=1e-2, precond="jacobi")
    assert np.all(res.relres <= 1e-8), res.relres
    for i, f in enumerate(freqs):
        o = wg_case.oracle_solution(float(f))
        assert np.abs(res.S[i] - o["S"]).max() <= 1e-6
    print(f"cocg jacobi iters={res.iters.tolist()}")


def test_sweep_cocg_aux_needs_fewer_iterations(sweep):
    """The gradient-space correction cuts the COCG iteration count on the
    waveguide (same systems, same tolerance)."""
    freqs = np.array([10e9])
    kw = dict(precision="fp64", tol=1e-8, fblock=1, restart=20000, max_cycles=30, method="cocg", inner_rel32=1e-2)
    ra = sweep.solve(freqs, precond="aux", **kw)
    rj = sweep.solve(freqs, precond="jacobi", **kw)
    assert np.all(ra.relres <= 1e-8) and np.all(rj.relres <= 1e-8)
    assert ra.iters.sum() * 3 < rj.iters.sum(), (ra.iters, rj.iters)
    assert np.abs(ra.S - rj.S).max() <= 1e-6


def test_vector_nodal_space_needs_fewer_iterations(sweep):
    """The vector-nodal functions in the auxiliary space (default) against
    the gradient space alone: fewer COCG iterations on the waveguide, the same
    S (both against SuperLU)."""
    freqs = np.array([8.5e9, 11.5e9])
    kw = dict(precision="fp64", tol=1e-8, fblock=2, restart=20000, max_cycles=30, method="cocg", inner_rel32=1e-2)
    try:
        sweep.set_vector_aux(False)
        rg = sweep.solve(freqs, **kw)
    finally:
        sweep.set_vector_aux(True)
    rv = sweep.solve(freqs, **kw)
    assert np.all(rg.relres <= 1e-8) and np.all(rv.relres <= 1e-8)
    print(f"iterations gradient space {rg.iters.ravel().tolist()}, with vector-nodal {rv.iters.ravel().tolist()}")
    assert rv.iters.sum() * 1.5 < rg.iters.sum(), (rv.iters, rg.iters)
    for i, f in enumerate(freqs):
        o = wg_case.oracle_solution(float(f))
        assert np.abs(rv.S[i] - o["S"]).max() <= 1e-6


def test_mixed_escalates_to_fp64_inner_solves(sweep):
    """With three refinement cycles the complex64 pass ends short of 1e-8;
    the unfinished systems continue with complex128 inner solves from their
    iterate and converge (escalation counted)."""
    freqs = np.array([8.5e9, 11.5e9])
    res = sweep.solve(freqs, precision="mixed", tol=1e-8, fblock=2, restart=20000, max_cycles=3,
                      method="cocg", inner_rel32=1e-2, fallback=False)
    assert res.fp64_escalated >= 1
    assert np.all(res.relres <= 1e-8), res.relres
    for i, f in enumerate(freqs):
        o = wg_case.oracle_solution(float(f))
        assert np.abs(res.S[i] - o["S"]).max() <= 1e-6
    print(f"escalated={res.fp64_escalated} iters={res.iters.tolist()}")


def test_uneven_meshes_solved_concurrently():
    """Independent jittered waveguides of uneven size, one device context
    each, solved from host threads at once: every point converges and the
    S-parameters equal the one-at-a-time solves bit for bit."""
    import threading
    from emerge_amd import workloads as w
    probs = w.waveguide_family(count=3, n_range=(6_000, 16_000), seed=3)
    sws = [FrequencySweep(p.mesh, p.basis, p.eps_r, p.inv_mu_r, w.waveguide_bcs(p)) for p in probs]
    freqs = np.array([9e9, 10e9, 11e9])
    kw = dict(precision="mixed", tol=1e-8, fblock=8, restart=20000, max_cycles=30, method="cocg",
              inner_rel32=1e-2, keep_fields=False)
    seq = [sw.solve(freqs, **kw) for sw in sws]
    out = [None] * len(sws)

    def run(i):
        out[i] = sws[i].solve(freqs, **kw)

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(sws))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for a, b in zip(seq, out):
        assert np.all(b.relres <= 1e-8)
        assert np.array_equal(a.S, b.S)
        # |S21| -> 1 on the lossless guide above cutoff
        assert np.all(np.abs(np.abs(b.S[:, 1, 0]) - 1) <= 2e-2)


def test_autotuned_block_width(sweep):
    """fblock='auto' times each candidate width on this mesh, picks one of
    them, and the tuned sweep still converges to the oracle's S."""
    freqs = np.linspace(8e9, 12e9, 12)
    f, lanes = sweep.autotune(freqs, candidates=(4, 8, 16))
    assert f in (4, 8, 16)
    assert lanes in sweep.spmv_lane_options(f, "mixed", False)
    tried = sweep.tuning["seconds_per_system_iteration"]
    assert {F for F, _ in tried} == {4, 8, 16}
    # every usable lane layout of every width was timed (F = 16 complex64:
    # 8 systems per lane, 2 lanes per entry, 8 to 32 entries in flight)
    assert {g for F, g in tried if F == 16} == set(sweep.spmv_lane_options(16, "mixed", False)) == {0, 8, 16, 32}
    res = sweep.solve(freqs[:3], precision="mixed", tol=1e-8, fblock="auto", restart=20000, max_cycles=30,
                      method="cocg", inner_rel32=1e-2)
    assert np.all(res.relres <= 1e-8)
    for i, fr in enumerate(freqs[:3]):
        o = wg_case.oracle_solution(float(fr))
        assert np.abs(res.S[i] - o["S"]).max() <= 1e-6
    print(f"autotuned fblock={f} timings={sweep.tuning}")


@pytest.mark.parametrize("safe32,nf", [(True, 2), (False, 2), (True, 4)])
def test_mixed_safe32_modes_match_superlu(sweep, safe32, nf):
    """Complex64 COCG with the FP64 real stiffness and FP64 accumulation
    (safe32) and without it: both converge to the oracle's S (nf = 4: a
    block of 8 systems, the wide-batch kernels of the FP64-accumulating
    path)."""
    freqs = np.linspace(8e9, 12e9, nf)
    res = sweep.solve(freqs, precision="mixed", tol=1e-8, fblock=2 * nf, restart=20000, max_cycles=30,
                      method="cocg", inner_rel32=1e-2, safe32=safe32)
    assert np.all(res.relres <= 1e-8), res.relres
    for i, f in enumerate(freqs):
        o = wg_case.oracle_solution(float(f))
        assert np.abs(res.S[i] - o["S"]).max() <= 1e-6
    print(f"safe32={safe32} iters={res.iters.tolist()} escalated={res.fp64_escalated}")


def test_auto_precision_policy(sweep):
    """precision='auto': complex64 inner solves on the coarse guide at
    X band, FP64 where the mesh is fine against the wavelength."""
    assert sweep.choose_precision([8e9, 12e9]) == "mixed"
    assert sweep.choose_precision([1e9]) == "fp64"
    o = wg_case.oracle_solution(F0)
    res = sweep.solve([F0], precision="auto", tol=1e-8)
    assert res.precision == "mixed"
    assert np.abs(res.S[0] - o["S"]).max() <= 1e-6
    with pytest.raises(ValueError):
        sweep.solve([F0], precision="fp16")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solver_plugin_with_attached_mesh(precision):
    """SolverHIP with the mesh attached (gradient-space correction, D_g =
    diag(G^T A G) formed on the device) on the systems SolveRoutine hands it
    (A[ix_(ids, ids)], E/solver.py:1266-1268): the lossy lumped-port cavity
    and the WR-90 guide against SuperLU, and fewer COCG iterations than
    without the mesh."""
    from tests import cavity_case
    cases = []
    o = cavity_case.oracle_solution(6e9)
    cc_ = cavity_case.case()
    cases.append(("cavity", cc_["mesh"], cc_["basis"], o, 1))
    w = wg_case.oracle_solution(F0)
    cw = wg_case.case()
    cases.append(("wr90", cw["mesh"], cw["basis"], w, 1))
    for name, mesh, basis, oc, port in cases:
        ids = oc["ids"]
        A = oc["K"].tocsc()[np.ix_(ids, ids)]
        b = (oc["b"] + oc["pv"][port])[ids]
        x_ref = oc["fields"][port][ids]
        from emerge_amd.solver import SimulationError
        plain = SolverHIP("", precision=precision, tol=1e-9, method="cocg", restart=20000, max_cycles=5)
        try:
            it0 = int(plain.solve(A, b, None, False, 0)[1].aux["iterations"])
        except SimulationError:   # Jacobi alone may not get there
            it0 = None
        att = SolverHIP("", precision=precision, tol=1e-9, method="cocg", restart=40000,
                        max_cycles=60).attach(mesh, basis, solve_ids=ids)
        x1, rep1 = att.solve(A, b, None, False, 0)
        it1 = int(rep1.aux["iterations"])
        rr = np.linalg.norm(b - A @ x1) / np.linalg.norm(b)
        dx = np.linalg.norm(x1 - x_ref) / np.linalg.norm(x_ref)
        print(f"{name} {precision}: iterations {it0} -> {it1} with the mesh, relres {rr:.1e}, |dx| {dx:.1e}",
              flush=True)
        assert rr <= 1e-9 and dx <= 1e-6
        assert rep1.aux["gradient_space"] == "True"
        assert it0 is None or it1 < it0
        # a second right-hand side on the same matrix reuses the device state
        x2, rep2 = att.solve(A, 2 * b, None, True, 1)
        assert np.linalg.norm(x2 - 2 * x1) <= 1e-6 * np.linalg.norm(x1)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_cocg_block_of_64(sweep, precision):
    """64 (frequency, port) systems in one COCG block: every residual <= 1e-8
    and S equal to the same systems solved in blocks of 32."""
    freqs = np.linspace(8e9, 12e9, 32)
    r64 = sweep.solve(freqs, precision=precision, fblock=64)
    r32 = sweep.solve(freqs, precision=precision, fblock=32)
    assert np.all(r64.relres <= 1e-8) and np.all(r32.relres <= 1e-8)
    assert np.abs(r64.S - r32.S).max() <= 1e-6
