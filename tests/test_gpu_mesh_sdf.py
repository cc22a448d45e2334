"""Mesh collider SDF built on the GPU (nb_mpm_set_collider_mesh) against the
oracle's buildMeshSDF (src/core/MeshSDF.h:85-126).

Both round every operation as written (no FMA contraction) and divide and
take square roots correctly rounded, so the SDFs must agree bit for bit, in
the reference's variant and in the exact one, for a box, a rotated box read
back from an STL file, triangle soups that span several LDS chunks and
several launches, and no triangles at all.  A run with the mesh installed must equal a run with the
same SDF handed over as a dense Morton array (setColliderSDF)."""
import numpy as np
import pytest

import oracle
from nebula_amd import mesh, mpm
from nebula_amd.mpm import MaterialModel
from nebula_amd.scenes import falling_box, make_batch_engine, make_engine

pytestmark = pytest.mark.gpu


def _scene():
    return falling_box("mesh", 40, 6000, 3e4, 0.3, MaterialModel.DRUCKER_PRAGER, np.random.default_rng(11))


def _meshes(tmp_path):
    box = mesh.box_mesh((0.1, 0.05, 0.12), (0.5, 0.2, 0.45))
    s = np.sqrt(1.0 / 3.0) * np.sin(0.4)  # 0.8 rad about (1, 1, 1)
    centred = mesh.box_mesh((-0.2, -0.075, -0.165), (0.2, 0.075, 0.165))
    rot = mesh.transform_triangles(centred, (0.3, 0.2, 0.3), (np.cos(0.4), s, s, s))
    mesh.save_binary_stl(tmp_path / "rot.stl", rot)
    rng = np.random.default_rng(2)
    v = rng.uniform(0.05, 0.6, (600, 3, 3)).astype(np.float32)
    nrm = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    soup = np.concatenate([nrm, v.reshape(600, 9)], axis=1).astype(np.float32)
    # small triangles spread over the domain: more than one launch's worth (2048)
    c = rng.uniform(0.05, 0.6, (4500, 1, 3))
    v2 = (c + rng.uniform(-0.02, 0.02, (4500, 3, 3))).astype(np.float32)
    n2 = np.cross(v2[:, 1] - v2[:, 0], v2[:, 2] - v2[:, 0])
    n2 /= np.linalg.norm(n2, axis=1, keepdims=True)
    soup2 = np.concatenate([n2, v2.reshape(4500, 9)], axis=1).astype(np.float32)
    return {"box": box, "rotated_stl": mesh.load_binary_stl(tmp_path / "rot.stl"), "soup_600": soup,
            "soup_4500": soup2, "empty": np.zeros((0, 12), np.float32)}


@pytest.mark.parametrize("exact", [False, True], ids=["reference", "exact"])
def test_mesh_sdf_matches_oracle_bitwise(gpu_lib, tmp_path, exact):
    sc = _scene()
    e = make_engine(sc)
    res = sc.cfg.gridRes()
    for name, tris in _meshes(tmp_path).items():
        print(f"mesh {name}: {tris.shape[0]} triangles, exact={exact}", flush=True)
        e.setColliderMesh(tris, exact=exact)
        got = e.getColliderSDF()
        want = oracle.mesh_sdf(tris, res, sc.cfg.cellSize, exact=exact)
        diff = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
        assert diff.size == 0, f"{name}: {diff.size} cells differ, first {diff[:5]}: {got[diff[:5]]} vs {want[diff[:5]]}"


def test_run_with_mesh_equals_run_with_dense_sdf(gpu_lib):
    sc = _scene()
    tris = mesh.box_mesh((0.1, 0.02, 0.1), (0.52, 0.12, 0.52))
    a, b = make_engine(sc), make_engine(sc)
    a.setColliderMesh(tris, exact=True)
    b.setColliderSDF(oracle.mesh_sdf(tris, sc.cfg.gridRes(), sc.cfg.cellSize, exact=True))
    for e in (a, b):
        e.substep(sc.dt, 40)
        e.sync()
    sa, sb = a.getState(), b.getState()
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
    assert np.isfinite(sa).all()


def test_mesh_sdf_refused_on_a_batch(gpu_lib):
    b = make_batch_engine([_scene(), _scene()])
    with pytest.raises(mpm.NebulaError):
        b.setColliderMesh(mesh.box_mesh((0.1, 0.1, 0.1), (0.2, 0.2, 0.2)))


def test_engine_with_collider_sdf_is_released_cleanly(gpu_lib):
    """Destroying an engine that holds a collider SDF frees each allocation
    once; a second free would leave an error that the next engine's launch
    checks report."""
    import gc
    sc = _scene()
    e = make_engine(sc)
    e.setColliderSphere(0.1, 0.3, 0.3, 0.3)
    del e
    gc.collect()
    f = make_engine(sc)
    f.setColliderSphere(0.1, 0.3, 0.3, 0.3)
    f.setColliderMesh(mesh.box_mesh((0.1, 0.1, 0.1), (0.2, 0.2, 0.2)))
