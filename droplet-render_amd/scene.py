"""Scene: the sparse octree and density bricks the render kernel reads.

Mirrors droplet-render's `Scene` (src/scene.h:146-159): `ob` is the node array
(OctreeStructure records), `pvol[VOLUME_BUFFER_FOG]` the bricks, `lvoxc` the
brick edge, `leafx` the brick counts. `Scene.from_grid` replaces the OpenVDB
part of Scene::Initialize (src/scene.cpp:840-917) for dense float grids.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class Scene:
    def __init__(self, ob, pvol_fog, lvoxc, mlevel=None, root_ce=None, voxel_size=None, pvol_sdf=None):
        self.ob = np.ascontiguousarray(ob)            # structured node array (see node_dtype)
        self.pvol = [None if pvol_sdf is None else np.ascontiguousarray(pvol_sdf, dtype=np.float32),
                     np.ascontiguousarray(pvol_fog, dtype=np.float32)]
        self.lvoxc = int(lvoxc)
        self.lvoxc3 = self.lvoxc ** 3
        self.leafx = [0 if pvol_sdf is None else int(self.pvol[0].size // self.lvoxc3),
                      int(self.pvol[1].size // self.lvoxc3)]
        self.index = len(self.ob) - 1
        self.mlevel = mlevel
        self.root_ce = root_ce
        self.voxel_size = voxel_size
        self._dev = None

    node_dtype = np.dtype([("ce", "<f4", (4,)), ("chn", "<u4", (8,)), ("volx", "<u4", (2,)), ("qval", "<f4", (2,))])

    @classmethod
    def from_grid(cls, fog, voxel_size=None, maxd=12, sdf=None):
        """Builds the octree from a dense (nz, ny, nx) float grid (x fastest);
        voxel (i,j,k) sits at world (i,j,k)*voxel_size (default 1/nx). `sdf`
        (same shape, world-unit signed distance, negative inside; `fog` may
        then be None) adds SDF leaves (drk_scene_build_dense)."""
        ref = fog if fog is not None else sdf
        nz, ny, nx = ref.shape
        fog = None if fog is None else np.ascontiguousarray(fog, dtype=np.float32)
        sdf = None if sdf is None else np.ascontiguousarray(sdf, dtype=np.float32)
        s = float(voxel_size) if voxel_size else 1.0 / nx
        out = C.POINTER(L.HostScene)()
        L.check(L.lib().drk_scene_build_dense(L.fptr(fog), L.fptr(sdf), nx, ny, nz, s, maxd, C.byref(out)))
        return cls._adopt(out, s)

    @classmethod
    def _adopt(cls, out, voxel_size=None):
        """Copies a drk_host_scene into a Scene and frees it."""
        try:
            hs = out.contents
            n = hs.node_count
            raw = np.ctypeslib.as_array(C.cast(hs.ob, C.POINTER(C.c_uint8)), shape=(n * 64,)).copy()
            ob = raw.view(cls.node_dtype)
            l3 = hs.lvoxc ** 3
            nf = hs.leafx[1] * l3
            pf = (np.ctypeslib.as_array(hs.pvol[1], shape=(nf,)).copy() if nf else np.zeros(0, np.float32))
            ns = hs.leafx[0] * l3
            ps = np.ctypeslib.as_array(hs.pvol[0], shape=(ns,)).copy() if ns else None
            sc = cls(ob, pf, hs.lvoxc, hs.mlevel, np.array(hs.root_ce[:], np.float32), voxel_size, ps)
        finally:
            L.lib().drk_host_scene_free(out)
        return sc

    def save(self, path):
        """Writes the node records and bricks to a scene blob (drk_scene_save)."""
        hs = L.HostScene()
        hs.ob = self.ob.ctypes.data_as(C.POINTER(L.OctreeNode))
        hs.node_count = self.node_count
        hs.pvol[0] = L.fptr(self.pvol[0]) if self.leafx[0] else None
        hs.pvol[1] = L.fptr(self.pvol[1]) if self.leafx[1] else None
        hs.leafx[0] = self.leafx[0]
        hs.leafx[1] = self.leafx[1]
        hs.lvoxc = self.lvoxc
        hs.mlevel = self.mlevel or 0
        rc = self.root_ce if self.root_ce is not None else np.zeros(4, np.float32)
        for i in range(4):
            hs.root_ce[i] = float(rc[i])
        L.check(L.lib().drk_scene_save(C.byref(hs), str(path).encode()))

    @classmethod
    def load(cls, path):
        """Reads a scene blob written by `save` (drk_scene_load)."""
        out = C.POINTER(L.HostScene)()
        L.check(L.lib().drk_scene_load(str(path).encode(), C.byref(out)))
        return cls._adopt(out)

    @property
    def node_count(self):
        return len(self.ob)

    def occupancy(self):
        """Fraction of leaves whose brick holds any density (leaf count over
        the full 8^(mlevel-1) leaf lattice)."""
        occ = int(np.count_nonzero(self.pvol[1].reshape(-1, self.lvoxc3).max(axis=1) > 0)) if self.leafx[1] else 0
        return occ / float(8 ** (self.mlevel - 1)) if self.mlevel else 0.0

    def device(self):
        """Uploads the scene once and returns the opaque drk_scene handle."""
        if self._dev is None:
            h = C.c_void_p()
            obp = self.ob.ctypes.data_as(C.POINTER(L.OctreeNode))
            sdf = self.pvol[0] if self.leafx[0] else None
            L.check(L.lib().drk_scene_create(obp, self.node_count, L.fptr(sdf), self.leafx[0], L.fptr(self.pvol[1]),
                                             self.leafx[1], self.lvoxc, C.byref(h)))
            self._dev = h
        return self._dev

    def footprint(self):
        """Device memory of the uploaded scene (nodes, child links, bricks), in bytes."""
        n = C.c_size_t()
        L.check(L.lib().drk_scene_footprint(self.device(), C.byref(n)))
        return int(n.value)

    def update_fog(self, pvol_fog, qval_fog=None):
        """Replaces the FOG bricks (host array or device pointer via torch) and optionally the majorants."""
        q = None if qval_fog is None else np.ascontiguousarray(qval_fog, dtype=np.float32)
        if hasattr(pvol_fog, "data_ptr"):
            ptr = C.cast(C.c_void_p(pvol_fog.data_ptr()), L._F)
        else:
            pvol_fog = np.ascontiguousarray(pvol_fog, dtype=np.float32)
            self.pvol[1] = pvol_fog
            ptr = L.fptr(pvol_fog)
        L.check(L.lib().drk_scene_update_fog(self.device(), ptr, L.fptr(q)))
        if q is not None:
            self.ob["qval"][:, L.DRK_VOLUME_FOG] = q

    def update_fog_device(self, pvol_fog, stream=None):
        """Replaces the FOG bricks from a CUDA tensor with a copy on `stream`
        (ordered with the stream's other work; no host synchronisation)."""
        assert pvol_fog.is_cuda and pvol_fog.dtype.itemsize == 4 and pvol_fog.is_contiguous()
        assert pvol_fog.numel() == self.leafx[1] * self.lvoxc3
        s = None if stream is None else C.c_void_p(stream.cuda_stream)
        L.check(L.lib().drk_scene_update_fog_device(self.device(), C.c_void_p(pvol_fog.data_ptr()), s))

    def set_walk(self, walk):
        """DRK_WALK_AUTO, DRK_WALK_OCTREE or DRK_WALK_GRID (next launches)."""
        L.check(L.lib().drk_scene_set_walk(self.device(), int(walk)))

    def walk_info(self):
        """(grid depth or 0, fraction of occupied grid cells, launches walk the grid)."""
        lv, occ, g = C.c_uint32(), C.c_float(), C.c_int()
        L.check(L.lib().drk_scene_walk_info(self.device(), C.byref(lv), C.byref(occ), C.byref(g)))
        return int(lv.value), float(occ.value), bool(g.value)

    def set_reference_fog(self, pvol_ref):
        """Reference FOG bricks for weighted tracking (host array or torch
        tensor of leafx*lvoxc^3 floats); None drops them."""
        if pvol_ref is None:
            L.check(L.lib().drk_scene_set_reference_fog(self.device(), None))
            return
        if hasattr(pvol_ref, "data_ptr"):
            assert pvol_ref.numel() == self.leafx[1] * self.lvoxc3
            ptr = C.c_void_p(pvol_ref.data_ptr())
        else:
            self._ref = np.ascontiguousarray(pvol_ref, dtype=np.float32).reshape(-1)
            assert self._ref.size == self.leafx[1] * self.lvoxc3
            ptr = C.c_void_p(self._ref.ctypes.data)
        L.check(L.lib().drk_scene_set_reference_fog(self.device(), ptr))

    def events_begin(self, samples, stream=None):
        """Opens the gradient window: the _device adjoint calls of this
        scene's kernels append their events for one reduction at
        events_end (`samples`: the pixel samples of all those calls)."""
        s = None if stream is None else C.c_void_p(stream.cuda_stream)
        L.check(L.lib().drk_scene_events_begin(self.device(), int(samples), s))

    def events_end(self, grad, stream=None):
        """Closes the gradient window: adds the reduction of its events to
        `grad` (a CUDA tensor of leafx*lvoxc^3 floats)."""
        assert grad.is_cuda and grad.dtype.itemsize == 4 and grad.numel() == self.leafx[1] * self.lvoxc3
        s = None if stream is None else C.c_void_p(stream.cuda_stream)
        L.check(L.lib().drk_scene_events_end(self.device(), C.c_void_p(grad.data_ptr()), s))

    def raise_majorants(self, factor=1.25, floor=None):
        """Multiplies every non-zero FOG majorant by `factor` (any value at or
        above the brick maximum keeps delta tracking unbiased); used to keep
        density perturbations and optimisation steps under fixed majorants.
        Leaves with majorant 0 (never resampled) stay transparent."""
        q = self.ob["qval"][:, L.DRK_VOLUME_FOG].copy()
        m = q > 0
        q[m] *= np.float32(factor)
        if floor is not None:
            q[m] = np.maximum(q[m], np.float32(floor))
        self.ob["qval"][:, L.DRK_VOLUME_FOG] = q
        if self._dev is not None:
            L.check(L.lib().drk_scene_update_fog(self._dev, L.fptr(self.pvol[1]), L.fptr(q)))

    def destroy(self):
        if self._dev is not None:
            L.lib().drk_scene_destroy(self._dev)
            self._dev = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
