"""Mesh topology built on the device (emg_topology_build) against the host
restatement of Mesh3D._pre_update (E/mesh3d.py:386-498), array by array and
bit for bit, on structured, jittered, graded and shuffled meshes."""
import numpy as np
import pytest

from emerge_amd import workloads as w
from emerge_amd.mesh import Mesh3D, kuhn_box

pytestmark = pytest.mark.gpu

FIELDS = ["edges", "tris", "tet_to_edge", "tet_to_tri", "tri_to_edge", "tri_to_tet", "edge_lengths"]


def _shuffled(seed):
    """Jittered box with vertex ids, tet order and each tet's local vertex
    order permuted."""
    nodes, tets = kuhn_box(5, 4, 3, size=(0.02, 0.015, 0.01), jitter=0.15, seed=seed)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(nodes.shape[1])
    inv = np.argsort(perm)
    nodes = nodes[:, perm]
    tets = inv[tets]
    tets = tets[:, rng.permutation(tets.shape[1])]
    for t in range(tets.shape[1]):
        tets[:, t] = tets[rng.permutation(4), t]
    return nodes, tets


def _cases():
    yield "cube", kuhn_box(6, 6, 6, jitter=0.15, seed=0)
    yield "shuffled", _shuffled(3)
    p = w.patch_antenna(h_slab=5e-3, n_sub=2, h_air=10e-3, inset=5e-3)
    yield "graded", (p.mesh.nodes, p.mesh.tets)


@pytest.mark.parametrize("name,mesh", list(_cases()), ids=lambda v: v if isinstance(v, str) else "")
def test_device_topology_is_bit_exact(hip_ctx, name, mesh):
    nodes, tets = mesh
    host = Mesh3D(nodes, tets)
    dev = Mesh3D.on_device(nodes, tets, hip_ctx)
    for f in FIELDS:
        a, b = getattr(host, f), getattr(dev, f)
        assert a.shape == b.shape, (name, f)
        assert np.array_equal(a, b), (name, f)
    # lookups work on the device-built object
    e = dev.get_edge(int(dev.edges[0, 5]), int(dev.edges[1, 5]))
    assert e == 5
    assert np.array_equal(dev.boundary_tris(), host.boundary_tris())


@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_dof_maps_and_pec_keep_match_host(jitter):
    """emg_dof_maps and emg_pec_keep on the device topology equal
    Nedelec2's maps and the complement of the PEC DOFs, bit for bit."""
    from emerge_amd._lib import Context
    from emerge_amd.basis import Nedelec2, pec_keep_ids
    from emerge_amd.mesh import Mesh3D, kuhn_box
    nodes, tets = kuhn_box(7, 4, 9, size=(0.02286, 0.01016, 0.04), jitter=jitter, seed=4)
    ctx = Context(0)
    md = Mesh3D.on_device(nodes, tets, ctx)
    fd = Nedelec2.on_device(md, ctx)
    mh = Mesh3D(nodes, tets)
    fh = Nedelec2(mh)
    for name in ("tet_to_field", "tri_to_field", "edge_to_field"):
        np.testing.assert_array_equal(getattr(fd, name), getattr(fh, name))
    assert fd.n_field == fh.n_field
    bt = mh.boundary_tris()
    pec = np.setdiff1d(bt, np.concatenate([mh.plane_tris(2, 0.0), mh.plane_tris(2, 0.04)]))
    removed = np.unique(np.concatenate([fh.edge_to_field[:, mh.tri_to_edge[:, pec].reshape(-1)].reshape(-1),
                                        fh.tri_to_field[:, pec].reshape(-1)]))
    keep = np.setdiff1d(np.arange(fh.n_field), removed)
    np.testing.assert_array_equal(pec_keep_ids(ctx, pec), keep)
    np.testing.assert_array_equal(pec_keep_ids(ctx, np.zeros(0, dtype=np.int64)), np.arange(fh.n_field))


@pytest.mark.parametrize("name,mesh", list(_cases()), ids=lambda v: v if isinstance(v, str) else "")
def test_device_topology_and_dof_maps_match_oracle(hip_ctx, name, mesh):
    """Device arrays against the set/dict oracle (oracle.topology, restating
    E/mesh3d.py:386-498 and E/elements/nedelec2.py:42-58) directly, and the
    device PEC solve_ids against the oracle's elimination
    (ASM/assembler.py:274-307, 406-407 as oracle.fem.assemble_freq_matrix)."""
    from emerge_amd.basis import Nedelec2, pec_keep_ids
    from oracle import topology as otopo
    nodes, tets = mesh
    print(f"topology vs oracle: {name} T={tets.shape[1]}", flush=True)
    ref = otopo.pre_update(nodes, tets)
    dev = Mesh3D.on_device(nodes, tets, hip_ctx)
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(dev, f), ref[f], err_msg=f"{name} {f}")
    fd = Nedelec2.on_device(dev, hip_ctx)
    maps = otopo.nedelec2_maps(ref, tets.shape[1])
    for key in ("tet_to_field", "tri_to_field", "edge_to_field"):
        np.testing.assert_array_equal(getattr(fd, key), maps[key], err_msg=f"{name} {key}")
    assert fd.n_field == maps["n_field"]
    # PEC on the boundary triangles whose centroid lies in the lower half
    bt = dev.boundary_tris()
    zc = nodes[2][dev.tris[:, bt]].mean(axis=0)
    pec = bt[zc < np.median(zc)]
    pec_ids = set()
    for e in ref["tri_to_edge"][:, pec].ravel():
        pec_ids.update(int(i) for i in maps["edge_to_field"][:, e])
    for t in pec:
        pec_ids.update(int(i) for i in maps["tri_to_field"][:, t])
    solve_ids = np.array([i for i in range(maps["n_field"]) if i not in pec_ids])
    np.testing.assert_array_equal(pec_keep_ids(hip_ctx, pec), solve_ids)
