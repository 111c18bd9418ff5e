"""Camera matrices in the reference's convention.

droplet-render builds its view with DirectXMath XMMatrixLookToRH and its
projection with XMMatrixPerspectiveRH from Blender's view frame, then passes
their inverses to RenderKernel::Initialize (src/main.cpp:71-115). Matrices are
row-major with the row-vector convention v' = v * M, so the translation sits
in row 3. These functions restate the published DirectXMath formulas.
"""
import math

import numpy as np


def _normalize(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def look_to_rh(eye, direction, up):
    """XMMatrixLookToRH: right-handed view matrix looking along `direction`."""
    eye = np.asarray(eye, dtype=np.float64)
    z = _normalize(-np.asarray(direction, dtype=np.float64))
    x = _normalize(np.cross(up, z))
    y = np.cross(z, x)
    m = np.eye(4)
    m[0, :3] = [x[0], y[0], z[0]]
    m[1, :3] = [x[1], y[1], z[1]]
    m[2, :3] = [x[2], y[2], z[2]]
    m[3, :3] = [-np.dot(x, eye), -np.dot(y, eye), -np.dot(z, eye)]
    return m


def perspective_fov_rh(fov_y, aspect, znear, zfar):
    """XMMatrixPerspectiveFovRH (vertical field of view in radians)."""
    h = 1.0 / math.tan(0.5 * fov_y)
    w = h / aspect
    r = zfar / (znear - zfar)
    m = np.zeros((4, 4))
    m[0, 0] = w
    m[1, 1] = h
    m[2, 2] = r
    m[2, 3] = -1.0
    m[3, 2] = r * znear
    return m


def camera(eye, target, fov_y_deg, w, h, up=(0.0, 0.0, 1.0), znear=0.1, zfar=100.0):
    """Returns (viewi, proji) as float32 4x4 arrays, ready for RenderKernel.

    Blender cameras use +z up; the scene frame of the configs follows it.
    """
    eye = np.asarray(eye, dtype=np.float64)
    d = np.asarray(target, dtype=np.float64) - eye
    view = look_to_rh(eye, d, np.asarray(up, dtype=np.float64))
    proj = perspective_fov_rh(math.radians(fov_y_deg), float(w) / float(h), znear, zfar)
    viewi = np.linalg.inv(view).astype(np.float32)
    proji = np.linalg.inv(proj).astype(np.float32)
    return np.ascontiguousarray(viewi), np.ascontiguousarray(proji)


def sun_direction(elevation_deg, azimuth_deg=0.0):
    """Unit direction toward the sun (+z up), as SunLight::direction holds it."""
    el, az = math.radians(elevation_deg), math.radians(azimuth_deg)
    return np.array([math.cos(el) * math.cos(az), math.cos(el) * math.sin(az), math.sin(el)], dtype=np.float32)
