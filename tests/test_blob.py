"""Scene blob: a built scene survives save/load bit for bit, and damaged or
foreign files are rejected through the C ABI with DRK_EINVAL."""
import numpy as np
import pytest


def _scene(pkg):
    g = pkg.configs.synth_spheres(48, 2, spheres=10)
    return pkg.Scene.from_grid(g)


def test_roundtrip_is_bit_exact(pkg, tmp_path):
    sc = _scene(pkg)
    p = tmp_path / "cloud.drk"
    sc.save(p)
    rt = pkg.Scene.load(p)
    assert rt.node_count == sc.node_count and rt.lvoxc == sc.lvoxc and rt.mlevel == sc.mlevel
    assert rt.ob.tobytes() == sc.ob.tobytes()
    assert np.array_equal(rt.pvol[1].view(np.uint32), sc.pvol[1].view(np.uint32))
    assert np.array_equal(rt.root_ce, sc.root_ce)
    assert p.stat().st_size == 64 + 64 * sc.node_count + 4 * sc.pvol[1].size


@pytest.mark.parametrize("damage", ["magic", "truncate", "flip", "trailing", "child"])
def test_damaged_files_are_rejected(pkg, tmp_path, damage):
    from droplet_render_amd import _lib as L
    sc = _scene(pkg)
    p = tmp_path / "cloud.drk"
    sc.save(p)
    b = bytearray(p.read_bytes())
    if damage == "magic":
        b[0] = ord("X")
    elif damage == "truncate":
        b = b[:-100]
    elif damage == "flip":
        b[64 + 64 * sc.node_count + 17] ^= 0x40        # a brick value: checksum mismatch
    elif damage == "trailing":
        b += b"\0"
    else:
        # a child index out of range, with a matching checksum
        ob = sc.ob.copy()
        ob["chn"][0, 0] = sc.node_count + 5
        bad = pkg.Scene(ob, sc.pvol[1], sc.lvoxc, sc.mlevel, sc.root_ce)
        bad.save(p)
        b = bytearray(p.read_bytes())
    p.write_bytes(bytes(b))
    with pytest.raises(L.DrkError) as e:
        pkg.Scene.load(p)
    assert e.value.code == L.DRK_EINVAL


def test_missing_file_is_an_error(pkg, tmp_path):
    from droplet_render_amd import _lib as L
    with pytest.raises(L.DrkError):
        pkg.Scene.load(tmp_path / "none.drk")


@pytest.mark.gpu
def test_loaded_scene_renders_the_same(pkg, gpu, tmp_path):
    wl = pkg.configs.tiny(n=32, w=32, h=24, spp=6, scattevs=4)
    p = tmp_path / "tiny.drk"
    wl.scene.save(p)
    k = wl.kernel()
    k.set_seed(3, 0)
    k.Render(0, 0, wl.w, wl.h, wl.spp)
    a = [x.copy() for x in k.phb]
    k.Destroy()
    wl.scene = pkg.Scene.load(p)
    k = wl.kernel()
    k.set_seed(3, 0)
    k.Render(0, 0, wl.w, wl.h, wl.spp)
    b = [x.copy() for x in k.phb]
    k.Destroy()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
