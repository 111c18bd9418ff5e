"""Frame driver: the Blender add-on's tile loop without Blender.

Restates `DropletRenderEngine.render` / `RenderTiles` / `RenderScene` /
`RenderShadow` (addon/__init__.py:87-207): the frame is covered by a grid of
tiles centred on it (edge tiles cropped), tiles are rendered nearest to the
frame centre first, each tile in passes of at most `samples_int` samples
until `samples_ext` are summed, and the passes are formed from the running
sums exactly as the add-on does (float64 numpy):

    COMBINED              = (cl + cs) * [1, 1, 1, 0.5] / dd
    TRANSMISSION_DIRECT   = cl.rgb / dd
    TRANSMISSION_INDIRECT = cs.rgb / dd
    SHADOW                = (Shadow pass of 100 samples).rgb / 100

with cl, cs the tile's `phb` sums and dd the samples so far.
"""
import math

import numpy as np


def tile_grid(width, height, tilew, tileh):
    """(x, y, w, h) of every tile, row-major (addon/__init__.py:176-201)."""
    nx = int(math.ceil(width / tilew))
    ny = int(math.ceil(height / tileh))
    out = []
    for y in range(ny):
        for x in range(nx):
            tx = x * tilew - 0.5 * (nx * tilew - width)
            ty = y * tileh - 0.5 * (ny * tileh - height)
            tw, th = tilew, tileh
            if tx + tilew > width:
                tw += int(width - (tx + tilew))
            if ty + tileh > height:
                th += int(height - (ty + tileh))
            tw += int(min(tx, 0))
            th += int(min(ty, 0))
            out.append((int(max(tx, 0)), int(max(ty, 0)), tw, th))
    return out


def tile_order(tiles, width, height, tilew, tileh):
    """RenderTiles' order: repeatedly the remaining tile whose nominal centre
    is nearest the frame centre (first in list order on ties)."""
    rest = list(tiles)
    out = []
    while rest:
        t = min(rest, key=lambda q: math.hypot(q[0] + 0.5 * tilew - 0.5 * width, q[1] + 0.5 * tileh - 0.5 * height))
        rest.remove(t)
        out.append(t)
    return out


def render_frame(kernel, width, height, tilew, tileh, samples_ext, samples_int, primary=True, shadow=False):
    """Renders a frame tile by tile; returns the passes as (height, width, c)
    float64 arrays. `kernel` was initialised for this image with a tile at
    least tilew x tileh."""
    passes = {}
    if primary:
        passes["COMBINED"] = np.zeros((height, width, 4))
        passes["TRANSMISSION_DIRECT"] = np.zeros((height, width, 3))
        passes["TRANSMISSION_INDIRECT"] = np.zeros((height, width, 3))
    if shadow:
        passes["SHADOW"] = np.zeros((height, width, 3))
    order = tile_order(tile_grid(width, height, tilew, tileh), width, height, tilew, tileh)
    if primary:
        sc = int(math.ceil(samples_ext / samples_int))
        for (x0, y0, tw, th) in order:
            cl = np.zeros((th, tw, 4))
            cs = cl.copy()
            dd = 0
            for i in range(sc):
                d1 = min(samples_ext - i * samples_int, samples_int)
                kernel.Render(x0, y0, tw, th, d1)
                ql, qe = kernel.phb
                dd += d1
                cl += ql
                cs += qe
            fd = 1.0 / float(dd)
            passes["COMBINED"][y0:y0 + th, x0:x0 + tw] = (cl + cs) * np.array([fd, fd, fd, 0.5 * fd])
            passes["TRANSMISSION_DIRECT"][y0:y0 + th, x0:x0 + tw] = cl[..., :3] * fd
            passes["TRANSMISSION_INDIRECT"][y0:y0 + th, x0:x0 + tw] = cs[..., :3] * fd
    if shadow:
        for (x0, y0, tw, th) in order:
            kernel.Shadow(x0, y0, tw, th, 100)
            cl = np.asarray(kernel.phb[0], np.float64)
            passes["SHADOW"][y0:y0 + th, x0:x0 + tw] = cl[..., :3] / 100.0
    return passes
