"""CPU checks of the instanced-scene host side: the generator's transforms
are rigid, the flattening restates the reference loader's baking
(scene_loader.ts:244-345: positions by the world matrix, normals and
tangents by the normal matrix, renormalised) and maps every flattened
triangle back to its (instance, mesh triangle)."""
import numpy as np


def _small():
    from dspt import scenes
    return scenes.instanced_scene(seed=7, n_meshes=4, n_instances=24, tri_range=(300, 3000), extent=6.0, height=3.0,
                                  width=32, height_px=32, spp=2, max_bounces=4, ibl_size=(32, 16))


def test_transforms_are_rigid():
    s = _small().scene
    R = s.o2w[:, :, :3].astype(np.float64)
    I = np.einsum("nij,nkj->nik", R, R)
    np.testing.assert_allclose(I, np.broadcast_to(np.eye(3), I.shape), atol=1e-6)
    np.testing.assert_allclose(np.linalg.det(R), 1.0, atol=1e-5)
    assert s.instanced_triangles == int(np.asarray(s.mesh_count)[s.inst_mesh].sum())


def test_flatten_bakes_world_space():
    s = _small().scene
    f, inst_of, mtri_of = s.flatten()
    assert f.num_triangles == s.instanced_triangles == len(inst_of) == len(mtri_of)
    ov = s.scene.vertexBuffer.reshape(-1, 20).astype(np.float64)
    otri = s.scene.triangleIndexBuffer.reshape(-1, 3)
    fv = f.vertexBuffer.reshape(-1, 20).astype(np.float64)
    ftri = f.triangleIndexBuffer.reshape(-1, 3)
    rng = np.random.default_rng(0)
    for k in rng.integers(0, len(inst_of), 200):
        i, mt = inst_of[k], mtri_of[k]
        M = s.o2w[i].astype(np.float64)
        for c in range(3):
            a, b = ov[otri[mt, c]], fv[ftri[k, c]]
            np.testing.assert_allclose(b[0:3], M[:, :3] @ a[0:3] + M[:, 3], atol=1e-5)
            np.testing.assert_allclose(b[4:7], M[:, :3] @ a[4:7], atol=1e-5)  # rigid: normal matrix == rotation
            assert b[3] == a[3]  # material index kept
            np.testing.assert_array_equal(b[8:12], a[8:12])
    # instance blocks in instance order, each the whole mesh
    np.testing.assert_array_equal(np.unique(inst_of), np.arange(s.num_instances))
    assert np.all(np.diff(inst_of) >= 0)


def test_torus_is_closed():
    from dspt.scenes import _torus
    p, n, uv, f = _torus(2000, 1.0, np.random.default_rng(1))
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert np.all(cnt == 2)
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-12)
