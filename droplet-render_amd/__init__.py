"""droplet-render's volumetric light-transport path, MI355X-native.

Package layout (the directory name has a dash, so it is loaded through
`load()` below as the module `droplet_render_amd`):

    _lib      ctypes binding of libdroplet_amd.so (include/droplet_render.h)
    scene     Scene: octree nodes + density bricks (src/scene.h:146-159)
    sampler   MiePhase / HGPhase / SunLight / HosekSky / NullEnv / MapEnv
    kernel    RenderKernel: Initialize / Render / Shadow / Destroy, phb
    camera    view / projection matrices in the reference's convention
    configs   the synthetic benchmark workloads
    build     in-tree hipcc build of the library
"""
from . import _lib, camera, sampler, scene, kernel, batch, configs, probe  # noqa: F401
from .scene import Scene  # noqa: F401
from .kernel import RenderKernel  # noqa: F401
from .sampler import MiePhase, HGPhase, SunLight, HosekSky, NullEnv, MapEnv  # noqa: F401
