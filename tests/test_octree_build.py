"""Octree build (src/scene.cpp:234-305, :676-762, :840-917): the library's C++
builder against the oracle's independent Python restatement (bit-exact), and
both against geometric facts that pin them."""
import numpy as np
import pytest

NO = 0xFFFFFFFF


def as_lists(sc):
    return [(tuple(np.float32(sc.ob["ce"][i])), tuple(int(x) for x in sc.ob["chn"][i]),
             tuple(int(x) for x in sc.ob["volx"][i]), tuple(np.float32(sc.ob["qval"][i]))) for i in range(len(sc.ob))]


def grids():
    rng = np.random.default_rng(11)
    g1 = np.zeros((24, 24, 24), np.float32)
    g1[5:9, 3:20, 11:13] = rng.random((4, 17, 2), dtype=np.float32)
    g2 = (rng.random((32, 32, 32)) > 0.985).astype(np.float32) * rng.random((32, 32, 32), dtype=np.float32)
    g3 = np.zeros((16, 40, 24), np.float32)
    g3[0, 0, 0] = 1.0
    g3[15, 39, 23] = 0.5
    g4 = np.zeros((40, 16, 24), np.float32)
    g4[5, 3, 9] = 1.0       # smaller than half a leaf: the depth formula goes below 0 (held at 1)
    return [("slab", g1), ("sparse", g2), ("corners", g3), ("one_voxel", g4)]


@pytest.mark.parametrize("name,grid", grids())
def test_cpp_build_matches_python_restatement(pkg, oracle, name, grid):
    sc = pkg.Scene.from_grid(grid)
    nodes, bricks, lvoxc, mlevel = oracle.build_octree(grid)
    assert (sc.lvoxc, sc.mlevel) == (lvoxc, mlevel)
    assert len(nodes) == sc.node_count
    for i, n in enumerate(nodes):
        assert np.array_equal(np.float32(n["ce"]), sc.ob["ce"][i]), (i, n["ce"], sc.ob["ce"][i])
        assert list(sc.ob["chn"][i]) == n["chn"], i
        assert list(sc.ob["volx"][i]) == n["volx"], i
        assert np.array_equal(np.float32(n["qval"]), sc.ob["qval"][i]), i
    assert np.array_equal(bricks.reshape(-1), sc.pvol[1])


def test_fbm_cloud_matches_python_restatement(pkg, oracle):
    g = pkg.configs.synth_fbm(32, 5, octaves=3, freq=3.0, threshold=0.2)
    sc = pkg.Scene.from_grid(g)
    nodes, bricks, _, _ = oracle.build_octree(g)
    assert len(nodes) == sc.node_count
    assert np.array_equal(bricks.reshape(-1), sc.pvol[1])


def test_maxd_doubles_brick_resolution(pkg, oracle):
    """Depth cap (src/scene.cpp:712-716): each level above maxd doubles lvoxc."""
    g = np.zeros((64, 64, 64), np.float32)
    g[::9, ::7, ::5] = 1.0
    full = pkg.Scene.from_grid(g)
    capped = pkg.Scene.from_grid(g, maxd=2)
    assert full.mlevel == 3 and full.lvoxc == 8
    assert capped.mlevel == 2 and capped.lvoxc == 16
    nodes, bricks, lvoxc, mlevel = oracle.build_octree(g, maxd=2)
    assert (lvoxc, mlevel) == (16, 2)
    assert np.array_equal(bricks.reshape(-1), capped.pvol[1])


@pytest.mark.parametrize("n,mlevel,nodes,leaves", [(64, 3, 73, 64), (128, 4, 585, 512)])
def test_dense_cube_counts(pkg, n, mlevel, nodes, leaves):
    """A fully active n^3 unit cube gives a complete octree: sum 8^l nodes,
    8^(mlevel-1) leaves of 16 voxels (leaf level is mlevel-1, src/scene.cpp:273)."""
    g = np.ones((n, n, n), np.float32)
    sc = pkg.Scene.from_grid(g)
    assert (sc.mlevel, sc.node_count, sc.leafx[1]) == (mlevel, nodes, leaves)
    leaf_e = sc.ob["ce"][sc.ob["volx"][:, 1] != NO][:, 3]
    assert np.allclose(leaf_e * 2 * n, 16.0)


def test_structural_invariants(pkg):
    g = pkg.configs.synth_spheres(64, 2, spheres=6)
    sc = pkg.Scene.from_grid(g)
    ob = sc.ob
    n = len(ob)
    level = np.full(n, -1)
    level[0] = 0
    first_touch = []
    for i in range(n):
        for c in ob["chn"][i]:
            if c:
                assert c > i                        # DFS creation order: children after parents
                level[c] = level[i] + 1
    leaves = np.nonzero(ob["volx"][:, 1] != NO)[0]
    assert set(level[leaves]) == {sc.mlevel - 1}     # leaves exactly at mlevel-1
    interior = np.nonzero(ob["volx"][:, 1] == NO)[0]
    assert np.all(level[interior] < sc.mlevel - 1)
    # volx is dense 0..leafx-1, assigned in node-creation (first touch) order
    assert np.array_equal(ob["volx"][leaves, 1], np.arange(len(leaves)))
    # majorant floor 0.01 and majorant = brick max, except the last node,
    # which the reference never resamples (its loop runs over [0, index)).
    bricks = sc.pvol[1].reshape(-1, sc.lvoxc3)
    for i in leaves:
        if i == n - 1:
            assert ob["qval"][i, 1] == 0.0 and not bricks[ob["volx"][i, 1]].any()
        else:
            assert ob["qval"][i, 1] == max(np.float32(0.01), bricks[ob["volx"][i, 1]].max())
    assert n - 1 in leaves


def test_leaves_cover_active_voxels(pkg, oracle):
    """Brute force: every active voxel lies in some leaf cube, and every leaf
    cube (grown by the 1.1 overlap margin) touches an active VDB-leaf box."""
    g = pkg.configs.synth_spheres(48, 4, spheres=5)
    sc = pkg.Scene.from_grid(g)
    s = 1.0 / 48
    ob = sc.ob
    leaves = ob[ob["volx"][:, 1] != NO]
    act = np.argwhere(g > 0)[:, ::-1] * s         # (x,y,z) world
    c, e = leaves["ce"][:, :3], leaves["ce"][:, 3:4]
    inside = np.all(np.abs(act[:, None, :] - c[None]) <= e[None] + 1e-6, axis=2).any(axis=1)
    assert inside.all()
    boxes = oracle.vdb_leaves(g)
    bc = np.array([[(mn[a] + mx[a]) * 0.5 * s for a in range(3)] for mn, mx in boxes])
    be = np.array([[0.5 * (mx[a] - mn[a] + 1) * s for a in range(3)] for mn, mx in boxes])
    for i in range(len(leaves)):
        ov = np.all(np.abs(bc - c[i]) <= be + 1.1 * e[i] + 1e-6, axis=1)
        assert ov.any(), i


def test_brick_vertices_sample_the_grid(pkg):
    """Brick vertex j sits at c - e*(7-2*idx)/7 (src/scene.cpp:892-895): with a
    grid that is linear in x the trilinear resampling reproduces the linear
    function at every vertex inside the grid."""
    n = 32
    x = np.arange(n, dtype=np.float32) / n
    g = np.broadcast_to(0.2 + 0.5 * x[None, None, :], (n, n, n)).astype(np.float32).copy()
    sc = pkg.Scene.from_grid(g)
    ob = sc.ob
    for i in np.nonzero(ob["volx"][:, 1] != NO)[0][:-1]:
        ce = ob["ce"][i]
        b = sc.pvol[1].reshape(-1, 8, 8, 8)[ob["volx"][i, 1]]
        vx = ce[0] - ce[3] * (7 - 2 * np.arange(8)) / 7.0
        ok = (vx >= 0) & (vx <= (n - 1) / n)
        expect = 0.2 + 0.5 * vx
        got = b[4, 4, :]
        assert np.allclose(got[ok], expect[ok], atol=2e-6), (i, got, expect)


def sphere_sdf(n, c=(0.5, 0.5, 0.5), r=0.25):
    """Signed distance (world units, negative inside) of a sphere on an n^3 grid, voxel (i,j,k) at (i,j,k)/n."""
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float64) / n
    return (np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r).astype(np.float32)


def test_sdf_build_matches_python_restatement(pkg, oracle):
    """SDF leaves (narrow band 4 voxels) before FOG leaves, SDF bricks and
    qval[SDF] = min distance, and fog leaves inside the surface dropped
    (src/scene.cpp:676-701, 874-904): C++ against the Python restatement."""
    n = 128
    sdf = sphere_sdf(n, r=0.42)
    fog = pkg.configs.synth_fbm(n, 5, octaves=3, freq=3.0, threshold=-1.0)   # fog everywhere
    sc = pkg.Scene.from_grid(fog, sdf=sdf)
    nodes, bricks, lvoxc, mlevel, sbricks = oracle.build_octree(fog, sdf=sdf)
    assert (sc.lvoxc, sc.mlevel) == (lvoxc, mlevel) and len(nodes) == sc.node_count
    for i, nd in enumerate(nodes):
        assert np.array_equal(np.float32(nd["ce"]), sc.ob["ce"][i]), i
        assert list(sc.ob["chn"][i]) == nd["chn"], i
        assert list(sc.ob["volx"][i]) == nd["volx"], i
        assert np.array_equal(np.float32(nd["qval"]), sc.ob["qval"][i]), i
    assert np.array_equal(bricks.reshape(-1), sc.pvol[1])
    assert np.array_equal(sbricks.reshape(-1), sc.pvol[0])
    # structure: some leaves carry both bricks, some fog leaves were dropped
    both = (sc.ob["volx"][:, 0] != NO) & (sc.ob["volx"][:, 1] != NO)
    assert sc.leafx[0] > 0 and both.any()
    assert (sc.ob["volx"][:, 1] != NO).sum() < sc.leafx[1]


def test_sdf_bricks_hold_the_clamped_distance(pkg):
    """At brick vertices that coincide with grid points the SDF brick holds the
    grid's distance held to the band; qval[SDF] is the brick minimum."""
    n = 32
    sdf = sphere_sdf(n, r=0.3)
    sc = pkg.Scene.from_grid(None, sdf=sdf)
    band = np.float32(4.0 / n)
    assert sc.leafx[1] == 0 and sc.leafx[0] > 0
    sb = sc.pvol[0].reshape(sc.leafx[0], -1)
    assert sb.min() >= -band and sb.max() <= band
    leaves = np.nonzero(sc.ob["volx"][:, 0] != NO)[0]
    for i in leaves[:-1]:
        if i == sc.node_count - 1:
            continue
        assert np.float32(sc.ob["qval"][i, 0]) == sb[sc.ob["volx"][i, 0]].min()
