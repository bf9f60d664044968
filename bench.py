#!/usr/bin/env python3
"""Benchmark of the EMerge FEM hot path on one MI355X.

Workload (one "step"): assembly of the Nedelec-2 curl-curl (E) and mass (B)
matrices of a 60^3 jittered Kuhn cube (1,296,000 tets, 8,316,360 complex
DOFs, per-tet complex eps_r, mu_r = 1) straight into CSR values on the device,
inputs resident in HBM.  The one-time pattern build is timed separately.

Prints one JSON line (rank 0).  ``--gpus`` above 1 is rejected: this build
runs on a single GPU.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "freq-points/s at ≤1e-8 rel. residual (mixed vs FP64 GPU); Mtets/s assembled"
HBM_PEAK_GBS = 8000.0


def assembly_algorithmic_bytes(T: int, V: int, nnz: int) -> int:
    """Compulsory bytes of one assembly: per tet 4 int32 vertex ids, an
    isotropic complex material pair (16 B as in the workload's eps_r with
    mu_r fixed), a 400-entry int32 scatter map; 3 doubles per vertex; complex128
    E and B values per CSR entry."""
    return T * (16 + 16 + 1600) + 24 * V + 32 * nnz


def cpu_baseline(n_tets_sample: int = 100_000) -> dict:
    """The reference's CPU path on a bounded sample of the same workload,
    timed on the host: the oracle's C++/OpenMP restatement of its per-tet
    kernel (tet-parallel, as the reference's numba prange; all host threads)
    plus scipy COO->CSR of E and B (single-threaded, as in the reference).
    Timed again with the element kernel on 4 threads, the reference's
    default NUMBA_NUM_THREADS (src/emerge/__init__.py:42-48)."""
    from oracle import curlcurl as cc
    from oracle import port
    from scipy.sparse import coo_matrix
    from emerge_amd.workloads import cube_assembly
    n = max(2, int(round((n_tets_sample / 6) ** (1 / 3))))
    prob = cube_assembly(n, seed=0)
    m, f = prob.mesh, prob.basis
    args = (m.nodes, m.tets, m.tris, m.edges, m.edge_lengths, f.tet_to_field, m.tet_to_edge,
            cc.isotropic(1.0 / prob.inv_mu_r), cc.isotropic(prob.eps_r))
    N = f.n_field

    def run(threads):
        port.set_threads(threads)
        port.matrix_builder(*args)  # warm (thread pool, page faults)
        t0 = time.perf_counter()
        dE, dB, r, c, nth = port.matrix_builder(*args)
        t1 = time.perf_counter()
        coo_matrix((dE, (r, c)), shape=(N, N)).tocsr()
        coo_matrix((dB, (r, c)), shape=(N, N)).tocsr()
        t2 = time.perf_counter()
        return int(nth), t1 - t0, t2 - t1

    nth, tk, tc = run(0)
    nth4, tk4, tc4 = run(4)
    port.set_threads(0)
    return {"value": m.n_tets / (tk + tc) / 1e6, "unit": "Mtets/s", "cores": nth, "kind": "port",
            "sample": f"{n}^3 jittered Kuhn cube ({m.n_tets} tets): C++/OpenMP port of the per-tet kernel "
                      f"({nth} threads) {tk:.2f} s + scipy COO->CSR (1 thread) {tc:.2f} s",
            "at_reference_threads": {"value": m.n_tets / (tk4 + tc4) / 1e6, "cores": nth4,
                                     "sample": f"same sample, per-tet kernel on {nth4} threads "
                                               f"(NUMBA_NUM_THREADS default) {tk4:.2f} s + COO->CSR {tc4:.2f} s"}}


def sweep_bench(n_plain: int, n_mixed: int, fblock: int, cpu_points: int, method: str = "cocg",
                idr_s: int = 4, restart: int = 20000, precond: str = "aux") -> dict:
    """WR-90 waveguide (Kuhn 11x5x18, 5,940 tets, 41,118 DOFs), TE10 ports,
    frequency points from the 201-point 8-12 GHz grid, both ports excited
    per point.  Plain FP64 GPU path: the FP64 solver one (frequency, port)
    system at a time.  Mixed path: complex64 inner iterations in FP64
    refinement, `fblock` systems per batch.  A point counts once every port
    reached ||b-Ax||/||b|| <= 1e-8 (FP64 residual on the device)."""
    from emerge_amd.sweep import FrequencySweep
    from emerge_amd.workloads import waveguide, waveguide_bcs
    prob = waveguide()
    bcs = waveguide_bcs(prob)
    sw = FrequencySweep(prob.mesh, prob.basis, prob.eps_r, prob.inv_mu_r, bcs)
    grid = np.linspace(8e9, 12e9, 201)
    plain_f = grid[np.linspace(0, 200, n_plain).astype(int)]
    mixed_f = grid[np.linspace(0, 200, n_mixed).astype(int)]
    kw = dict(method=method, idr_s=idr_s, restart=restart, max_cycles=400 if method == "gmres" else 40,
              keep_fields=False, inner_rel32={"gmres": 1e-5, "idr": 1e-4, "cocg": 1e-2}[method], precond=precond)
    kw_m = dict(kw)
    if method == "idr":  # IDR(s) runs in FP64 only; the mixed path stays GMRES-IR
        kw_m.update(method="gmres", restart=100, max_cycles=400, inner_rel32=1e-5)
    print(f"[bench] sweep: warm-up", file=sys.stderr, flush=True)
    sw.solve(plain_f[:1], precision="fp64", fblock=1, **kw)  # warm-up (code objects, allocations)
    print(f"[bench] sweep: fp64 plain, {len(plain_f)} points", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    rp = sw.solve(plain_f, precision="fp64", fblock=1, **kw)
    tp = time.perf_counter() - t0
    tuning = None
    lanes = 0
    fb_mixed = fblock
    if fblock == "auto":
        fblock, lanes = sw.autotune(mixed_f, precision="mixed", method=kw_m["method"], precond=precond)
        tuning = {f"F{k[0]}_gr{k[1]}": round(v * 1e6, 3) for k, v in sw.tuning["seconds_per_system_iteration"].items()}
        print(f"[bench] sweep: autotuned fblock {fblock} (us per system-iteration {tuning})", file=sys.stderr,
              flush=True)
    print(f"[bench] sweep: fp64 plain {tp:.2f} s; mixed batched, {len(mixed_f)} points", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    if tuning is not None:
        fb_mixed = "auto"  # the tuned width and lane layout (cached)
    rm = sw.solve(mixed_f, precision="mixed", fblock=fb_mixed, **kw_m)
    tm = time.perf_counter() - t0
    print(f"[bench] sweep: mixed {tm:.2f} s; fp64 batched at the same block width", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    rb = sw.solve(mixed_f, precision="fp64", fblock=fblock, **kw)
    tb = time.perf_counter() - t0
    print(f"[bench] sweep: fp64 batched {tb:.2f} s", file=sys.stderr, flush=True)
    ok_p = int(np.sum(np.all(rp.relres <= 1e-8, axis=1)))
    ok_m = int(np.sum(np.all(rm.relres <= 1e-8, axis=1)))
    ok_b = int(np.sum(np.all(rb.relres <= 1e-8, axis=1)))
    # |dS| against the FP64 direct solve (oracle + SuperLU fixture of all 201
    # points, tests/golden/sweeps.npz), where the points are on the grid
    dS = {}
    fx = os.path.join(ROOT, "tests", "golden", "sweeps.npz")
    if os.path.exists(fx):
        G = np.load(fx, allow_pickle=False)
        gi = {float(f): i for i, f in enumerate(G["wr90_freqs"])}
        for name, r, fr in (("mixed_batched", rm, mixed_f), ("fp64_batched", rb, mixed_f), ("fp64_plain", rp, plain_f)):
            idx = [gi.get(float(f)) for f in fr]
            if all(i is not None for i in idx):
                dS[name] = float(np.abs(r.S - G["wr90_S"][idx]).max())
    out = {
        "workload": "WR-90 22.86x10.16x40 mm, Kuhn 11x5x18 (5,940 tets, 41,118 DOFs), TE10 ports, 8-12 GHz",
        "krylov": {"idr": f"IDR({idr_s})", "gmres": f"GMRES({restart})",
                   "cocg": "COCG on (A+A^T)/2 in FP64 refinement"}[method],
        "preconditioner": ("Jacobi + auxiliary space: nodal gradients and vector-nodal functions with diagonal solves" if method == "cocg" and precond == "aux"
                           else "Jacobi"),
        "fp64_plain": {"points": len(plain_f), "converged": ok_p, "seconds": round(tp, 3),
                       "freq_points_per_s": round(ok_p / tp, 4),
                       "mean_iters_per_system": float(rp.iters.mean()), "max_relres": float(rp.relres.max())},
        "mixed_batched": {"points": len(mixed_f), "converged": ok_m, "seconds": round(tm, 3),
                          "freq_points_per_s": round(ok_m / tm, 4), "fblock": fblock,
                          "fblock_tuning_us_per_system_iteration": tuning,
                          "spmv_entries_in_flight": lanes,
                          "fp64_escalated_systems": int(rm.fp64_escalated),
                          "mean_iters_per_system": float(rm.iters.mean()), "max_relres": float(rm.relres.max()),
                          "max_dS_vs_direct": dS.get("mixed_batched")},
        "fp64_batched": {"points": len(mixed_f), "converged": ok_b, "seconds": round(tb, 3),
                         "freq_points_per_s": round(ok_b / tb, 4), "fblock": fblock,
                         "mean_iters_per_system": float(rb.iters.mean()), "max_relres": float(rb.relres.max()),
                         "max_dS_vs_direct": dS.get("fp64_batched")},
    }
    out["fp64_plain"]["max_dS_vs_direct"] = dS.get("fp64_plain")
    fp = max(out["fp64_plain"]["freq_points_per_s"], 1e-30)
    fb = max(out["fp64_batched"]["freq_points_per_s"], 1e-30)
    out["speedup_mixed_vs_fp64"] = round(out["mixed_batched"]["freq_points_per_s"] / fp, 3)
    out["speedup_mixed_vs_fp64_batched"] = round(out["mixed_batched"]["freq_points_per_s"] / fb, 3)
    out["speedup_fp64_batched_vs_plain"] = round(fb / fp, 3)
    if cpu_points > 0:
        from oracle import curlcurl as cc, fem
        from oracle import ports as oports
        from emerge_amd.workloads import WR90_A, WR90_B
        m, f = prob.mesh, prob.basis
        E, B = cc.tet_mass_stiffness_matrices(m, f.tet_to_field, cc.isotropic(prob.eps_r),
                                              cc.isotropic(1.0 / prob.inv_mu_r))
        L = prob.meta["length"]
        o_ports = [oports.RectangularWaveguide(bcs[1].tri_ids, 1, oports.CS([1, 0, 0], [0, 1, 0], [0, 0, 1],
                                               [WR90_A / 2, WR90_B / 2, 0.0]), (WR90_A, WR90_B)),
                   oports.RectangularWaveguide(bcs[2].tri_ids, 2, oports.CS([1, 0, 0], [0, 1, 0], [0, 0, 1],
                                               [WR90_A / 2, WR90_B / 2, L]), (WR90_A, WR90_B))]
        maps = dict(edge_to_field=f.edge_to_field, tri_to_field=f.tri_to_field)
        t0 = time.perf_counter()
        for fr in grid[np.linspace(0, 200, cpu_points).astype(int)]:
            K, b, pv, ids = fem.assemble_freq_matrix(m, maps, E, B, bcs[0].tri_ids, o_ports, fr)
            fem.solve_all_ports(K, b, pv, ids)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"freq_points_per_s": round(cpu_points / tc, 4), "cores": 1, "kind": "port",
                               "sample": f"{cpu_points} points: oracle per-frequency assembly + scipy SuperLU "
                                         f"(reference options), {tc:.1f} s"}
    return out


def config_bench(name: str, nf: int, fblock: int, precision: str = "fp64") -> dict:
    """Full-size solves of the other named workloads (one JSON object):

    * cavity: PEC cavity 60x40x30 mm, Kuhn 48x32x24 (221,184 tets), dielectric
      block, probe-fed lumped port, 5 GHz: mixed (complex64 inner, FP64
      refinement) against FP64 inner solves -- |dS11| and field difference;
    * patch: slab antenna in an air box with a first-order ABC (999,600
      tets), `nf` of the 51 points over 2-3 GHz (spread over the band),
      FP64 inner solves in blocks of `fblock`;
    * family: 16 jittered WR-90 sections of 20k-120k DOFs, 16 points each
      over 8-12 GHz, one context per mesh, solved from 16 host threads at
      once and one after another."""
    import threading
    from emerge_amd import workloads as w
    from emerge_amd.sweep import FrequencySweep
    kw = dict(method="cocg", restart=40000, max_cycles=40, inner_rel32=1e-2, keep_fields=True)
    out = {"config": name}
    if name == "cavity":
        p = w.cavity()
        sw = FrequencySweep(p.mesh, p.basis, p.eps_r, p.inv_mu_r, w.cavity_bcs(p))
        res = {}
        # one-time set-up (auxiliary space on the host and device, symmetric
        # part, code objects) outside the timed solves: a short solve of each
        for prec in ("mixed", "fp64", "m32"):
            sw.solve([5e9], precision=prec, fblock=1, fallback=False, **dict(kw, restart=50, max_cycles=1))
        for prec in ("mixed", "fp64", "m32"):
            t0 = time.perf_counter()
            r = sw.solve([5e9], precision=prec, fblock=1, **kw)
            res[prec] = (r, time.perf_counter() - t0)
            print(f"[bench] cavity {prec} {res[prec][1]:.2f} s relres {r.relres.max():.2e}", file=sys.stderr,
                  flush=True)
        (rm, tm), (rf, tf) = res["mixed"], res["fp64"]
        a, b = rm.fields[(0, 1)], rf.fields[(0, 1)]
        out.update(workload="cavity 60x40x30 mm, Kuhn 48x32x24 (221,184 tets), 5 GHz, lumped port",
                   dofs=int(p.basis.n_field), reduced=int(sw.n_red),
                   mixed={"seconds": round(tm, 3), "relres": float(rm.relres.max()), "iters": int(rm.iters.sum()),
                          "escalated": rm.fp64_escalated, "S11": [float(rm.S[0, 0, 0].real), float(rm.S[0, 0, 0].imag)]},
                   fp64={"seconds": round(tf, 3), "relres": float(rf.relres.max()), "iters": int(rf.iters.sum()),
                         "S11": [float(rf.S[0, 0, 0].real), float(rf.S[0, 0, 0].imag)]},
                   m32={"seconds": round(res["m32"][1], 3), "relres": float(res["m32"][0].relres.max()),
                        "iters": int(res["m32"][0].iters.sum()),
                        "dS11_vs_fp64": float(abs(res["m32"][0].S[0, 0, 0] - rf.S[0, 0, 0]))},
                   dS11=float(abs(rm.S[0, 0, 0] - rf.S[0, 0, 0])),
                   field_rel_diff=float(np.linalg.norm(a - b) / np.linalg.norm(b)))
    elif name == "patch":
        t_set = time.perf_counter()
        p = w.patch_antenna()
        print(f"[bench] patch: mesh of {p.mesh.n_tets} tets built", file=sys.stderr, flush=True)
        sw = FrequencySweep(p.mesh, p.basis, p.eps_r, p.inv_mu_r, w.patch_bcs(p))
        print(f"[bench] patch: matrices assembled ({time.perf_counter() - t_set:.1f} s of set-up)", file=sys.stderr,
              flush=True)
        grid = np.linspace(2e9, 3e9, 51)
        freqs = grid[np.round(np.linspace(0, 50, nf)).astype(int)]
        kw["keep_fields"] = False
        # a report per block of systems (a full sweep runs for minutes)
        t0 = time.perf_counter()
        sw.progress = lambda done, total: print(f"[bench] patch: {done}/{total} systems solved "
                                                f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
        r = sw.solve(freqs, precision=precision, fblock=fblock, **kw)
        dt = time.perf_counter() - t0
        ok = int(np.sum(np.all(r.relres <= 1e-8, axis=1)))
        out.update(workload="patch antenna slab 60x60x1.52 mm in an air box, first-order ABC, graded Kuhn "
                            f"70x70x34 ({p.mesh.n_tets} tets)", dofs=int(p.basis.n_field), reduced=int(sw.n_red),
                   precision=precision, points=len(freqs), converged=ok, seconds=round(dt, 2),
                   escalated=int(r.fp64_escalated),
                   freq_points_per_s=round(ok / dt, 4), iters=r.iters.ravel().tolist(),
                   relres_max=float(r.relres.max()),
                   s11_db={f"{f / 1e9:.2f}": round(float(20 * np.log10(abs(r.S[i, 0, 0]))), 3)
                           for i, f in enumerate(freqs)})
    elif name == "modal":
        # the WR-90 sweep of the bench line with modal ports: the TE10 mode
        # of each port face from the boundary-mode analysis at 8 GHz
        from emerge_amd.modal import ModalPort, modal_analysis
        from emerge_amd import bc as hbc
        p = w.waveguide()
        ana = w.waveguide_bcs(p)
        pec = [b for b in ana if isinstance(b, hbc.PEC)]
        rw = [b for b in ana if getattr(b, "is_port", False)]
        ports = [ModalPort(b.tri_ids, b.port_number, b.cs) for b in rw]
        from emerge_amd._lib import Context
        mctx = Context(0)
        t0 = time.perf_counter()
        for pt in ports:
            modal_analysis(mctx, p.mesh, pt, p.eps_r, p.inv_mu_r, pec, 8e9)
        t_modal = time.perf_counter() - t0
        mctx.close()
        freqs = np.linspace(8e9, 12e9, nf)
        kws = dict(kw, keep_fields=False, fblock=64, restart=20000)
        swm = FrequencySweep(p.mesh, p.basis, p.eps_r, p.inv_mu_r, pec + ports)
        swa = FrequencySweep(p.mesh, p.basis, p.eps_r, p.inv_mu_r, ana)
        swm.solve(freqs[:2], precision="mixed", **dict(kws, restart=50, max_cycles=1, fallback=False))
        swa.solve(freqs[:2], precision="mixed", **dict(kws, restart=50, max_cycles=1, fallback=False))
        t0 = time.perf_counter()
        rm = swm.solve(freqs, precision="mixed", **kws)
        tm = time.perf_counter() - t0
        t0 = time.perf_counter()
        ra = swa.solve(freqs, precision="mixed", **kws)
        ta = time.perf_counter() - t0
        ok = int(np.sum(np.all(rm.relres <= 1e-8, axis=1)))
        out.update(workload="WR-90 Kuhn 11x5x18 between two modal ports (boundary-mode analysis at 8 GHz), "
                            f"{nf} points over 8-12 GHz, mixed batched F = 64",
                   modal_analysis_s=round(t_modal, 3), beta_8GHz=[float(np.real(pt.get_beta(2 * np.pi * 8e9 / 299792458)))
                                                                    for pt in ports],
                   points=nf, converged=ok, seconds=round(tm, 3), freq_points_per_s=round(ok / tm, 2),
                   analytic_ports_freq_points_per_s=round(nf / ta, 2),
                   max_dS_vs_analytic_ports=float(np.abs(rm.S - ra.S).max()))
    elif name in ("family", "family_union"):
        # family_union: only the one-system leg with per-mesh scalars (for
        # kernel traces of that leg alone)
        probs = w.waveguide_family()
        freqs = np.linspace(8e9, 12e9, nf)
        kf = dict(kw, keep_fields=False, fblock=32, restart=20000)
        seq = None
        if name == "family":
            sws = [FrequencySweep(q.mesh, q.basis, q.eps_r, q.inv_mu_r, w.waveguide_bcs(q)) for q in probs]
            for sw in sws[:2]:
                sw.solve(freqs[:2], precision="mixed", **kf)
            t0 = time.perf_counter()
            seq = [sw.solve(freqs, precision="mixed", **kf) for sw in sws]
            ts = time.perf_counter() - t0
            conc = [None] * len(sws)

            def run(i):
                conc[i] = sws[i].solve(freqs, precision="mixed", **kf)

            th = [threading.Thread(target=run, args=(i,)) for i in range(len(sws))]
            t0 = time.perf_counter()
            for t in th:
                t.start()
            for t in th:
                t.join()
            tc = time.perf_counter() - t0
            ok_s = sum(int(np.sum(np.all(r.relres <= 1e-8, axis=1))) for r in seq)
            ok_c = sum(int(np.sum(np.all(r.relres <= 1e-8, axis=1))) for r in conc)
            for sw in sws:
                sw.ctx.close()
            out.update(sequential={"seconds": round(ts, 3), "converged": ok_s, "freq_points_per_s": round(ok_s / ts, 3)},
                       concurrent={"seconds": round(tc, 3), "converged": ok_c, "freq_points_per_s": round(ok_c / tc, 3)},
                       max_dS_seq_vs_conc=float(max(np.abs(a.S - b.S).max() for a, b in zip(seq, conc))),
                       mean_iters_per_system_single=float(np.mean([r.iters.mean() for r in seq])))
        # the same meshes as one block-diagonal system: one launch per kernel
        # for all meshes, with per-mesh Krylov scalars (row segments) and,
        # for comparison, with scalars global over the union
        from emerge_amd.family import FamilySweep
        for segd in ((True, False) if name == "family" else (True,)):
            fam = FamilySweep(probs, [w.waveguide_bcs(q) for q in probs], segmented=segd)
            fam.solve(freqs[:2], precision="mixed", **kf)
            t0 = time.perf_counter()
            Sf, bound, rf = fam.solve(freqs, precision="mixed", **kf)
            tf = time.perf_counter() - t0
            ok_f = int(np.sum(np.all(bound <= 1e-8, axis=1))) * len(probs)
            leg = {"seconds": round(tf, 3), "converged_points": ok_f, "freq_points_per_s": round(ok_f / tf, 3),
                   "reduced_dofs": int(fam.sweep.n_red), "max_iters_per_system": int(rf.iters.max()),
                   "mean_iters_per_system": float(rf.iters.mean())}
            if seq is not None:
                leg["max_dS_vs_sequential"] = float(max(np.abs(Sf[m] - seq[m].S).max() for m in range(len(probs))))
            out["one_system" if segd else "one_system_global_scalars"] = leg
            fam.ctx.close()
        out.update(workload="16 jittered WR-90 sections (Kuhn 11x5xnz), DOFs U[20k,120k] seed 0",
                   dofs=[int(q.basis.n_field) for q in probs], points_per_mesh=len(freqs))
    elif name == "farfield":
        out.update(farfield_bench())
    elif name == "element":
        out.update(element_bench())
    else:
        raise ValueError(name)
    return out


def element_bench(n: int = 60, reps: int = 10) -> dict:
    """Batched dense element matrices (emg_element_matrices) of the n^3
    jittered cube, kept on the device: the vector-unit kernel (compile-time
    sparse tables) against the FP64 matrix-core kernel (dense tables).
    Algorithmic bytes per tet: 16 connectivity + 32 material + 12,800
    written (K and M, 400 complex128 each) + 24 per vertex."""
    from emerge_amd._lib import Context
    from emerge_amd.workloads import cube_assembly
    prob = cube_assembly(n, seed=0)
    m, f = prob.mesh, prob.basis
    ctx = Context(0)
    ctx.mesh_set(m.nodes, m.tets, f.tet_to_field, f.n_field)
    ctx.material_set(prob.eps_r, prob.inv_mu_r)
    T = m.n_tets
    algo = T * (16 + 32 + 12800) + 24 * m.n_nodes
    res = {"workload": f"element matrices of the {n}^3 jittered Kuhn cube ({T} tets), device-resident output",
           "algorithmic_bytes": algo}
    for mode, name in ((0, "vector"), (1, "mfma")):
        for _ in range(2):
            ctx.call("emg_element_matrices", mode, None, None)
        ctx.synchronize()
        ctx.timer_start()
        for _ in range(reps):
            ctx.call("emg_element_matrices", mode, None, None)
        ms = ctx.timer_stop() / reps
        res[name] = {"ms": round(ms, 4), "Mtets_per_s": round(T / ms / 1e3, 2),
                     "GB_per_s": round(algo / ms / 1e6, 1), "hbm_frac": round(algo / ms / 1e6 / HBM_PEAK_GBS, 4)}
        print(f"[bench] element {name}: {ms:.3f} ms", file=sys.stderr, flush=True)
    ctx.close()
    return res


def farfield_bench(n: int = 60, reps: int = 5) -> dict:
    """Stratton-Chu far field over the surface of the n^3 Kuhn cube (edge-
    centre samples, synthetic complex E and H from PCG64 seed 0, all above
    the 1e-3 level) at 1 degree over the whole sphere (181 x 361
    directions), k0 = 2 pi / (side / 4); before it, E and curl E of a random
    solution sampled at the same points on the cube's mesh (point grid built
    by the first call).  Rate in (direction, sample) pairs
    per second, wall time per call (host copies included); CPU baseline:
    the oracle's NumPy restatement on a bounded set of directions."""
    from emerge_amd._lib import Context, ptr
    from emerge_amd.basis import Nedelec2
    from emerge_amd.farfield import FarFieldSurface, stratton_chu
    from emerge_amd.mesh import Mesh3D, kuhn_box
    nodes, tets = kuhn_box(n, n, n, size=(1.0, 1.0, 1.0))
    m = Mesh3D(nodes - 0.5, tets)
    surf = FarFieldSurface(m, m.boundary_tris(), origin=(0.0, 0.0, 0.0))
    M = surf.edge_centers.shape[1]
    rng = np.random.default_rng(0)
    # field sampling at the surface samples (emg_field_points) from a random
    # solution vector on the same mesh: first call builds the point grid
    f = Nedelec2(m)
    fctx = Context(0)
    fctx.mesh_set(m.nodes, m.tets, f.tet_to_field, f.n_field)
    xs = rng.standard_normal(f.n_field) + 1j * rng.standard_normal(f.n_field)
    pts = np.ascontiguousarray(surf.edge_centers)
    Es = np.empty((3, M), dtype=np.complex128)
    Cs = np.empty((3, M), dtype=np.complex128)
    own = np.empty(M, dtype=np.int32)
    samp = []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        fctx.call("emg_field_points", ptr(xs), M, ptr(pts), ptr(Es), ptr(Cs), ptr(own))
        samp.append(time.perf_counter() - t0)
    E = (1 + rng.random((3, M))) * np.exp(2j * np.pi * rng.random((3, M)))
    H = (1 + rng.random((3, M))) * np.exp(2j * np.pi * rng.random((3, M))) / 376.73031366857
    T, P = np.meshgrid(np.radians(np.arange(181.0)), np.radians(np.arange(361.0)), indexing="ij")
    T, P = T.ravel(), P.ravel()
    k0 = 2 * np.pi / 0.25
    ctx = Context(0)
    Eo, _, _ = stratton_chu(E, H, surf, T, P, k0, ctx=ctx)
    t0 = time.perf_counter()
    for _ in range(reps):
        Eo, _, _ = stratton_chu(E, H, surf, T, P, k0, ctx=ctx)
    dt = (time.perf_counter() - t0) / reps
    pairs = M * T.shape[0]
    res = {"workload": f"far field: surface of the {n}^3 Kuhn unit cube ({M} edge-centre samples), "
                       f"{T.shape[0]} directions (1 degree), synthetic fields",
           "samples": M, "directions": int(T.shape[0]), "ms_per_call": round(dt * 1e3, 3),
           "pairs_per_s": pairs / dt, "flops_per_pair": 53,
           "note": "wall time per call, host copies included; kernel time in the rocprof summary",
           "field_sampling": {"tets": int(m.n_tets), "points": M, "unowned": int((own < 0).sum()),
                              "first_call_ms": round(samp[0] * 1e3, 2),
                              "ms_per_call": round(float(np.mean(samp[1:])) * 1e3, 3)}}
    if not os.environ.get("EMG_NO_CPU_BASELINE"):
        from oracle import farfield as off
        sel = np.arange(0, T.shape[0], 131)
        t0 = time.perf_counter()
        Er, _ = off.stratton_chu_ff(E, H, surf.edge_centers, surf.weighted_normals, T[sel], P[sel], k0)
        tc = time.perf_counter() - t0
        res["max_rel_diff_vs_oracle"] = float(np.abs(Eo[:, sel] - Er).max() / np.abs(Er).max())
        res["cpu_baseline"] = {"pairs_per_s": M * sel.shape[0] / tc, "cores": 1, "kind": "port",
                               "sample": f"{sel.shape[0]} directions x {M} samples, oracle NumPy "
                                         f"restatement, {tc:.1f} s"}
        res["speedup_vs_cpu"] = round(res["pairs_per_s"] / res["cpu_baseline"]["pairs_per_s"], 1)
    return res


def hbm_probe(nbytes: int = 11_514_094_384, reps: int = 10) -> dict:
    """Achievable HBM rates for plain streams as large as the 60^3 cube's
    E and B values (11.51 GB): torch's fill (write only), copy (half read,
    half write) and sum (read only) kernels, timed with events."""
    import torch
    dev = torch.device("cuda:0")

    def timed(f):
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    n = nbytes // 8
    x = torch.empty(n, dtype=torch.float64, device=dev)
    t_fill = timed(lambda: x.fill_(1.0))
    t_sum = timed(lambda: torch.sum(x))
    del x
    y = torch.ones(n // 2, dtype=torch.float64, device=dev)
    z = torch.empty_like(y)
    t_copy = timed(lambda: z.copy_(y))
    gb = n * 8 / 1e9
    return {"bytes": n * 8, "fill_ms": round(t_fill, 4), "fill_TBps": round(gb / t_fill, 3),
            "copy_ms": round(t_copy, 4), "copy_TBps": round(gb / t_copy, 3),
            "sum_ms": round(t_sum, 4), "sum_TBps": round(gb / t_sum, 3)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=60, help="cube cells per side")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep-plain", type=int, default=4, help="FP64 plain-path frequency points")
    ap.add_argument("--sweep-mixed", type=int, default=201, help="mixed batched frequency points")
    ap.add_argument("--fblock", default="64", help="systems per batch (1..64, 64 with COCG only) or 'auto' "
                                                   "(tuned per mesh)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--sweep-only", action="store_true", help="profile helper: run only the solve sweep")
    ap.add_argument("--method", default="cocg", choices=["gmres", "idr", "cocg"])
    ap.add_argument("--idr-s", type=int, default=4)
    ap.add_argument("--restart", type=int, default=None,
                    help="GMRES dimension / IDR matvecs per cycle / COCG iterations per cycle")
    ap.add_argument("--config", choices=["cavity", "patch", "family", "family_union", "farfield", "element", "modal"], default=None,
                    help="run one full-size solve workload instead of the default bench")
    ap.add_argument("--config-points", type=int, default=None, help="frequency points of --config")
    ap.add_argument("--config-precision", default="fp64", choices=["fp64", "mixed", "m32"],
                    help="inner-solve precision of --config patch")
    ap.add_argument("--precond", default="aux", choices=["aux", "jacobi"],
                    help="COCG preconditioner: Jacobi + auxiliary gradient space, or Jacobi alone")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="counter calibration: only run 3 device-to-device copies of this many MiB")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="only time plain device fill / copy / sum streams of the assembly's output size "
                         "(the achievable HBM write and read rates)")
    args = ap.parse_args()
    args.fblock = args.fblock if args.fblock == "auto" else int(args.fblock)
    if args.restart is None:
        args.restart = {"gmres": 100, "idr": 30000, "cocg": 20000}[args.method]
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}); the multi-GPU sharding is "
              "documented, not launched", file=sys.stderr)
        return 2

    if args.copy_probe_mib > 0:
        # A plain copy of known size: its FETCH_SIZE / WRITE_SIZE readings
        # calibrate the counters used for the assembly kernel's HBM traffic.
        import torch
        nb = args.copy_probe_mib << 20
        a = torch.ones(nb, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nb}), flush=True)
        return 0

    if args.hbm_probe:
        print(json.dumps(hbm_probe()), flush=True)
        return 0

    from emerge_amd._lib import Context
    from emerge_amd.assembler import MeshOnDevice
    from emerge_amd.workloads import cube_assembly

    if args.config:
        nf = args.config_points or {"cavity": 1, "patch": 8, "family": 16, "family_union": 16, "farfield": 0, "element": 0,
                                    "modal": 201}[args.config]
        fb = 8 if args.fblock == "auto" else args.fblock
        print(json.dumps(config_bench(args.config, nf, fb if args.config == "patch" else 32, args.config_precision)),
              flush=True)
        return 0

    if args.sweep_only:
        print(json.dumps({"sweep": sweep_bench(args.sweep_plain, args.sweep_mixed, args.fblock, 0, args.method,
                                               args.idr_s, args.restart, args.precond)}), flush=True)
        return 0

    t0 = time.perf_counter()
    prob = cube_assembly(args.n, seed=0)
    m, f = prob.mesh, prob.basis
    t_mesh = time.perf_counter() - t0
    print(f"[bench] mesh T={m.n_tets} V={m.n_nodes} N={f.n_field} ({t_mesh:.1f} s)", file=sys.stderr, flush=True)

    ctx = Context(0)
    # mesh topology (E/mesh3d.py:386-498): host NumPy restatement vs the
    # device build (checked bit-exact here as well)
    from emerge_amd.mesh import Mesh3D
    t0 = time.perf_counter()
    mh = Mesh3D(m.nodes, m.tets)
    t_topo_host = time.perf_counter() - t0
    Mesh3D.on_device(m.nodes, m.tets, ctx)  # warm (code objects, allocations)
    t0 = time.perf_counter()
    md = Mesh3D.on_device(m.nodes, m.tets, ctx)
    t_topo_dev = time.perf_counter() - t0
    topo_exact = all(np.array_equal(getattr(mh, k), getattr(md, k))
                     for k in ("edges", "tris", "tet_to_edge", "tet_to_tri", "tri_to_edge", "tri_to_tet",
                               "edge_lengths"))
    del mh, md
    print(f"[bench] topology host {t_topo_host:.2f} s, device {t_topo_dev:.3f} s (bit-exact: {topo_exact})",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    dm = MeshOnDevice(ctx, m.nodes, m.tets, f.tet_to_field, f.n_field)
    ctx.synchronize()
    t_pattern = time.perf_counter() - t0
    stats = ctx.pattern_stats()
    print(f"[bench] pattern nnz={dm.nnz} chunks={stats['chunks']} ({t_pattern:.1f} s)", file=sys.stderr,
          flush=True)
    ctx.material_set(prob.eps_r, prob.inv_mu_r)

    for _ in range(args.warmup):
        ctx.assemble()
    ctx.synchronize()
    ctx.timer_start()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.assemble()
    ms_total = ctx.timer_stop()
    wall = time.perf_counter() - t0
    ms = ms_total / args.steps
    print(f"[bench] assembly {ms:.3f} ms/step (events), wall {wall / args.steps * 1e3:.3f} ms/step",
          file=sys.stderr, flush=True)

    T, V, nnz = m.n_tets, m.n_nodes, dm.nnz
    algo = assembly_algorithmic_bytes(T, V, nnz)
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on
    # this workload (copy-calibrated); reported only for the 60^3 cube it was
    # taken on and only while the kernel source is the one it was measured on
    traffic = None
    tf = os.path.join(ROOT, "profiles", "assembly_cube60_hbm_traffic.json")
    if args.n == 60 and os.path.exists(tf):
        with open(tf) as fh:
            rec = json.load(fh)
        with open(os.path.join(ROOT, rec["source"]), "rb") as fh:
            same = hashlib.sha256(fh.read()).hexdigest() == rec["source_sha256"]
        traffic = int(rec["traffic_bytes"]) if same else None
    achieved = algo / (ms * 1e-3) / 1e9
    value = T / (ms * 1e-3) / 1e6

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline()
        print(f"[bench] cpu baseline {cpu['value']:.4f} Mtets/s", file=sys.stderr, flush=True)

    sweep = None
    if not args.no_sweep:
        sweep = sweep_bench(args.sweep_plain, args.sweep_mixed, args.fblock,
                            0 if args.no_cpu_baseline else 2, args.method, args.idr_s, args.restart, args.precond)
        print(f"[bench] sweep {json.dumps(sweep)}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mtets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"assembly: unit cube {args.n}^3 hexes x 6 Kuhn tets, interior jitter 0.15h (seed 0), "
                        "eps_r=U[1,10](1-j U[0,0.02]), mu_r=1; FP64 Nedelec-2 E and B into CSR",
            "tets": T, "dofs": f.n_field, "nnz": nnz,
            "pattern_build_s": round(t_pattern, 3),
            "topology_host_s": round(t_topo_host, 3), "topology_device_s": round(t_topo_dev, 4),
            "topology_bit_exact": bool(topo_exact),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
        "sweep": sweep,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
