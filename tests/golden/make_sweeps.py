"""Writes tests/golden/sweeps.npz: S-parameter fixtures of the solve
workloads from the oracle (oracle assembly + SuperLU with the reference's
options + the oracle's S extraction, MW/microwave_3d.py:1167-1303), for GPU
tests and the bench's |dS| report:

* wr90_S: the WR-90 40 mm guide (Kuhn 11 x 5 x 18, TE10 ports) at all 201
  points of the 8-12 GHz grid (np.linspace(8e9, 12e9, 201));
* family_S: the 16 jittered WR-90 sections of waveguide_family() (seed 0,
  20k-120k DOFs) at 9 and 11 GHz each;
* patch_S: the reduced patch antenna of tests/patch_case.py at its two
  frequencies.

Run from the repository root: python tests/golden/make_sweeps.py [workers]
(about 15 minutes on 8 cores; one BLAS thread per worker; finished points
are kept in $TMPDIR/emg_sweeps_cache so an interrupted run resumes).
"""
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

WR90_GRID = np.linspace(8e9, 12e9, 201)
FAMILY_FREQS = np.array([9e9, 11e9])


CACHE = os.path.join(os.environ.get("TMPDIR", "/tmp"), "emg_sweeps_cache")


def _cached(name, fn):
    path = os.path.join(CACHE, name + ".npy")
    if os.path.exists(path):
        return np.load(path)
    out = np.asarray(fn())
    np.save(path + ".tmp.npy", out)
    os.replace(path + ".tmp.npy", path)
    return out


def _init():
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)


def _wr90(freq):
    def run():
        from tests import wg_case
        return wg_case.oracle_solution(float(freq))["S"]
    return _cached(f"wr90_{freq:.6e}", run)


_family_cache = {}


def _family(args):
    i, freq = args
    return _cached(f"family_{i}_{freq:.6e}", lambda: _family_run(i, freq))


def _family_run(i, freq):
    from emerge_amd import workloads as w
    from oracle import curlcurl as cc
    from oracle import fem
    from oracle import ports as oports
    if i not in _family_cache:
        p = w.waveguide_family()[i]
        bcs = w.waveguide_bcs(p)
        m, f = p.mesh, p.basis
        E, B = cc.tet_mass_stiffness_matrices(m, f.tet_to_field, cc.isotropic(p.eps_r),
                                              cc.isotropic(1.0 / p.inv_mu_r))
        a, b, L = p.meta["a"], p.meta["b"], p.meta["length"]
        o_ports = [oports.RectangularWaveguide(bcs[k + 1].tri_ids, k + 1,
                                               oports.CS([1, 0, 0], [0, 1, 0], [0, 0, 1], [a / 2, b / 2, z]),
                                               (a, b)) for k, z in enumerate((0.0, L))]
        _family_cache.clear()
        _family_cache[i] = (m, f, E, B, bcs[0].tri_ids, o_ports)
    m, f, E, B, pec, o_ports = _family_cache[i]
    maps = dict(edge_to_field=f.edge_to_field, tri_to_field=f.tri_to_field)
    K, bb, pv, ids = fem.assemble_freq_matrix(m, maps, E, B, pec, o_ports, float(freq))
    fields = fem.solve_all_ports(K, bb, pv, ids)
    return fem.s_parameters(m, f.tet_to_field, fields, o_ports, 2 * np.pi * freq / fem.C0)


def _patch(freq):
    def run():
        from tests import patch_case
        return patch_case.oracle_S(float(freq))
    return _cached(f"patch_{freq:.6e}", run)


def main():
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 6
    t0 = time.time()
    from tests import patch_case
    os.makedirs(CACHE, exist_ok=True)
    with Pool(workers, initializer=_init) as pool:
        fam_jobs = [(i, fr) for i in range(16) for fr in FAMILY_FREQS]
        r_fam = pool.map_async(_family, fam_jobs, chunksize=2)
        r_patch = pool.map_async(_patch, list(patch_case.FREQS))
        r_wr = pool.map_async(_wr90, list(WR90_GRID), chunksize=8)
        wr = np.array(r_wr.get())
        print(f"wr90 done ({time.time() - t0:.0f} s)", flush=True)
        fam = np.array(r_fam.get()).reshape(16, len(FAMILY_FREQS), 2, 2)
        print(f"family done ({time.time() - t0:.0f} s)", flush=True)
        patch = np.array(r_patch.get())
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sweeps.npz")
    np.savez_compressed(out, wr90_freqs=WR90_GRID, wr90_S=wr, family_freqs=FAMILY_FREQS, family_S=fam,
                        patch_freqs=np.array(patch_case.FREQS), patch_S=patch)
    print("wrote", out, f"({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
