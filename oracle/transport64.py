"""Float64 light-transport expectations (TEST INFRASTRUCTURE ONLY).

Independent of the oracle's tracker: the density field is the FP64 trilinear
interpolation of the scene's bricks (fp64.trilinear), transmittance is a
numerical integral of that field along a ray, and the expected pixel values
of a few estimator settings are quadratures over those transmittances. The
tests compare the oracle's (and the GPU's) Monte Carlo images with them, so a
misreading of the reference's estimator that the oracle and the kernels share
still fails a test.

What the expectations restate from the reference (src/kernel.cpp):
  * camera rays collide with density sigma_e * rho; the collision density
    along the ray is sigma_e rho(t) exp(-sigma_e int_0^t rho) (:274-315);
  * a secondary ray starting inside a leaf restarts free flight at that
    leaf's entry point (behind its origin): fog-only leaves are always
    entered at tra, even tra < td (:263-269), so its escape probability is
    exp(-sigma_e * int from the leaf entry to the exit);
  * the last-created leaf is never resampled and has majorant 0: it is
    transparent (scene.cpp:870, kernel.cpp:274);
  * light is added only when a ray with r > 0 escapes (:325); env light is
    env(rd)^2.2 (:362-364); the shadow ray's environment term is dropped
    (sky2 = 0, :409);
  * the MIS weights are w1 = p1/(p1 + Pdf(srd)) and w2 = p3/(p3 + Pdf(lrd))
    with the sun's Pdf 1/(2 pi (1 - cos^2 alpha)) inside its cone and 0
    outside (KernelSampler.cpp:326-330), while lrd is sampled uniformly in
    the cone (:332-345), and both contributions are scaled by sigma_s/sigma_e
    once (:416-419).
"""
import numpy as np

from . import fp64

NB = 0xFFFFFFFF


class Field:
    """FP64 density field of a Scene's FOG bricks: point location by descent
    of the octree (child slot x + 2y + 4z, bit set on the upper half), the
    trilinear lookup of the leaf's brick, zero outside leaves and in the
    last-created node (never resampled: majorant 0, transparent)."""

    def __init__(self, scene):
        ob = scene.ob
        self.ce = ob["ce"].astype(np.float64)
        self.chn = ob["chn"].astype(np.int64)
        self.volx = ob["volx"].astype(np.int64)
        self.leaf = (self.volx[:, 0] != NB) | (self.volx[:, 1] != NB)
        self.last = len(ob) - 1
        self.lvoxc = scene.lvoxc
        self.bricks = scene.pvol[1].reshape(-1, scene.lvoxc3)
        r = self.ce[0]
        self.lo, self.hi = r[:3] - r[3], r[:3] + r[3]

    def locate(self, p):
        """Leaf node containing each point (-1: none)."""
        p = np.asarray(p, np.float64).reshape(-1, 3)
        node = np.zeros(len(p), np.int64)
        inside = np.all((p >= self.lo) & (p <= self.hi), axis=1)
        node[~inside] = -1
        for _ in range(32):
            act = node >= 0
            act[act] = ~self.leaf[node[act]]
            if not act.any():
                break
            c = self.ce[node[act]]
            q = p[act]
            slot = (q[:, 0] >= c[:, 0]).astype(np.int64) + 2 * (q[:, 1] >= c[:, 1]) + 4 * (q[:, 2] >= c[:, 2])
            ch = self.chn[node[act], slot]
            node[act] = np.where(ch == 0, -1, ch)
        return node

    def density(self, p):
        """fp64.trilinear of each point's leaf brick, gathered for all points
        at once (tests check it against fp64.trilinear leaf by leaf)."""
        p = np.asarray(p, np.float64).reshape(-1, 3)
        node = self.locate(p)
        out = np.zeros(len(p))
        ok = (node >= 0) & (node != self.last)
        ok[ok] = self.volx[node[ok], 1] != NB
        if not ok.any():
            return out
        nd = node[ok]
        q = p[ok]
        ce = self.ce[nd]
        n = self.lvoxc
        ni = 0.5 * (q - (ce[:, :3] - ce[:, 3:4])) * (n - 1) / ce[:, 3:4]
        fl = np.floor(ni)
        a = np.clip(fl, 0, n - 1).astype(np.int64)
        c = np.clip(np.ceil(ni), 0, n - 1).astype(np.int64)
        f = ni - np.maximum(fl, 0)
        base = self.volx[nd, 1] * n ** 3
        flat = self.bricks.reshape(-1)
        acc = np.zeros(len(q))
        for dz in (0, 1):
            z = c[:, 2] if dz else a[:, 2]
            wz = f[:, 2] if dz else 1 - f[:, 2]
            for dy in (0, 1):
                y = c[:, 1] if dy else a[:, 1]
                wy = wz * (f[:, 1] if dy else 1 - f[:, 1])
                for dx in (0, 1):
                    x = c[:, 0] if dx else a[:, 0]
                    acc += wy * (f[:, 0] if dx else 1 - f[:, 0]) * flat[base + (z * n + y) * n + x]
        out[ok] = acc
        return out

    def density_by_leaf(self, p):
        """The same field evaluated leaf by leaf with fp64.trilinear."""
        p = np.asarray(p, np.float64).reshape(-1, 3)
        node = self.locate(p)
        out = np.zeros(len(p))
        ok = (node >= 0) & (node != self.last)
        ok[ok] = self.volx[node[ok], 1] != NB
        for nd in np.unique(node[ok]):
            m = node == nd
            out[m] = fp64.trilinear(self.bricks[self.volx[nd, 1]], p[m], self.ce[nd], self.lvoxc)
        return out

    def exit_t(self, x, w):
        """Parameter where rays x + t w leave the root box (>= 0)."""
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = np.where(w > 0, (self.hi - x) / w, np.where(w < 0, (self.lo - x) / w, np.inf))
        return np.maximum(np.min(t1, axis=1), 0.0)

    def entry_t(self, x, w):
        """Parameter (<= 0) of the entry point, along w, of the leaf that
        contains x; 0 when x lies in no leaf."""
        x = np.asarray(x, np.float64).reshape(-1, 3)
        w = np.asarray(w, np.float64).reshape(-1, 3)
        node = self.locate(x)
        t = np.zeros(len(x))
        m = node >= 0
        c = self.ce[node[m]]
        lo, hi = c[:, :3] - c[:, 3:4], c[:, :3] + c[:, 3:4]
        with np.errstate(divide="ignore", invalid="ignore"):
            ta = (lo - x[m]) / w[m]
            tb = (hi - x[m]) / w[m]
        tmin = np.where(w[m] == 0, -np.inf, np.minimum(ta, tb))
        t[m] = np.minimum(np.max(tmin, axis=1), 0.0)
        return t

    def tau(self, x, w, a, b, steps=192, chunk=1 << 20):
        """Midpoint integral of the density over [a, b] along x + t w, per ray."""
        x = np.asarray(x, np.float64).reshape(-1, 3)
        w = np.asarray(w, np.float64).reshape(-1, 3)
        a = np.broadcast_to(np.asarray(a, np.float64), (len(x),))
        b = np.broadcast_to(np.asarray(b, np.float64), (len(x),))
        out = np.zeros(len(x))
        per = max(1, chunk // steps)
        f = (np.arange(steps) + 0.5) / steps
        for i in range(0, len(x), per):
            s = slice(i, i + per)
            ln = np.maximum(b[s] - a[s], 0.0)
            t = a[s, None] + f[None, :] * ln[:, None]
            pts = x[s, None, :] + t[..., None] * w[s, None, :]
            rho = self.density(pts.reshape(-1, 3)).reshape(t.shape)
            out[s] = rho.sum(axis=1) * ln / steps
        return out

    def transmittance_secondary(self, x, w, sigma_e, steps=192):
        """Escape probability of a secondary ray from x along w: free flight
        restarts at the entry of the leaf containing x."""
        x = np.asarray(x, np.float64).reshape(-1, 3)
        w = np.asarray(w, np.float64).reshape(-1, 3)
        a = self.entry_t(x, w)
        b = self.exit_t(x, w)
        return np.exp(-sigma_e * self.tau(x, w, a, b, steps))

    def camera_collisions(self, ro, rd, sigma_e, cells=256):
        """Cells of a camera ray through the root box: midpoints (n, 3) and
        the probability that the ray's first collision lies in each cell
        (density taken constant per cell; the probabilities add up to
        1 - exp(-sigma_e tau) exactly)."""
        ro = np.asarray(ro, np.float64)
        rd = np.asarray(rd, np.float64)
        rd = rd / np.linalg.norm(rd)
        with np.errstate(divide="ignore", invalid="ignore"):
            ta = (self.lo - ro) / rd
            tb = (self.hi - ro) / rd
        t0 = np.max(np.where(rd == 0, -np.inf, np.minimum(ta, tb)))
        t1 = np.min(np.where(rd == 0, np.inf, np.maximum(ta, tb)))
        t0 = max(t0, 0.0)
        if not t1 > t0:
            return np.zeros((0, 3)), np.zeros(0), rd
        h = (t1 - t0) / cells
        t = t0 + (np.arange(cells) + 0.5) * h
        pts = ro[None, :] + t[:, None] * rd[None, :]
        rho = self.density(pts)
        cum = np.concatenate([[0.0], np.cumsum(rho * h)])[:-1]
        prob = np.exp(-sigma_e * cum) * (1.0 - np.exp(-sigma_e * rho * h))
        return pts, prob, rd


def basis(d):
    """Two unit vectors orthogonal to d and to each other."""
    d = np.asarray(d, np.float64) / np.linalg.norm(d)
    a = np.array([1.0, 0.0, 0.0]) if abs(d[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    b1 = np.cross(d, a)
    b1 /= np.linalg.norm(b1)
    return b1, np.cross(d, b1)


def cone_quadrature(d, cos_alpha, n_mu=6, n_phi=16):
    """Directions and solid-angle weights of a Gauss-Legendre (cos) x
    uniform (azimuth) rule over the cone of half-angle alpha around d."""
    x, wx = np.polynomial.legendre.leggauss(n_mu)
    mu = cos_alpha + (x + 1.0) * 0.5 * (1.0 - cos_alpha)
    wmu = wx * 0.5 * (1.0 - cos_alpha)
    phi = 2.0 * np.pi * (np.arange(n_phi) + 0.5) / n_phi
    b1, b2 = basis(d)
    d = np.asarray(d, np.float64) / np.linalg.norm(d)
    M, P = np.meshgrid(mu, phi, indexing="ij")
    W = np.repeat(wmu[:, None], n_phi, axis=1) * (2.0 * np.pi / n_phi)
    st = np.sqrt(np.maximum(1.0 - M * M, 0.0))
    dirs = M[..., None] * d + st[..., None] * (np.cos(P)[..., None] * b1 + np.sin(P)[..., None] * b2)
    return dirs.reshape(-1, 3), W.reshape(-1)


def sphere_quadrature(d, n_mu=32, n_phi=32):
    """Directions and weights of a Gauss-Legendre x uniform rule over the
    whole sphere, with the polar axis along d."""
    return cone_quadrature(d, -1.0, n_mu, n_phi)


def hg(g, ct):
    return (1.0 - g * g) / (4.0 * np.pi * (1.0 + g * g - 2.0 * g * np.asarray(ct, np.float64)) ** 1.5)


def sun_pdf_reference(cos_alpha):
    """SunLight::Pdf inside the cone, as the reference computes it
    (KernelSampler.cpp:326-330): 1/(2 pi (1 - cos^2 alpha))."""
    return 1.0 / (2.0 * np.pi * (1.0 - cos_alpha * cos_alpha))


class SingleScatter:
    """Expected per-camera-sample values of one pixel with scattevs = 1 and
    an HG phase function (sampled exactly, by inversion), from the
    reference's recursion: after the camera ray's collision at x, the shadow
    ray toward the sun cone carries albedo * w2 * T(x, lrd) * sun, and the
    phase-sampled continuation carries albedo * w1 * T(x, srd) * (sun if srd
    lies in the cone, plus env); the shadow ray's environment term is
    dropped. The transmittances are computed once; `values` combines them for
    a given sun pdf, albedo factor and shadow-ray environment term, so that
    tests can also show that plausible misreadings give other images."""

    def __init__(self, field, ro, rd, sigma_s, sigma_a, g, sun_dir, cos_alpha, cells=128, n_mu=6, n_phi=16,
                 sphere=(24, 24), steps=160):
        self.sig_e = sigma_s + sigma_a
        self.alb = sigma_s / self.sig_e
        self.cos_alpha = cos_alpha
        pts, prob, d = field.camera_collisions(ro, rd, self.sig_e, cells)
        keep = prob > 0
        self.prob = prob[keep]
        pts = pts[keep]
        self.omega = 2.0 * np.pi * (1.0 - cos_alpha)   # normaliser of the uniform cone sampling
        cd, self.cw = cone_quadrature(sun_dir, cos_alpha, n_mu, n_phi)
        self.ph_c = hg(g, cd @ d)
        sd, self.sw = sphere_quadrature(d, *sphere)
        self.ph_s = hg(g, sd @ d)
        n = len(pts)
        if n == 0:
            self.Tc = np.zeros((0, len(cd)))
            self.Ts = np.zeros((0, len(sd)))
            return
        self.Tc = field.transmittance_secondary(np.repeat(pts, len(cd), axis=0), np.tile(cd, (n, 1)), self.sig_e,
                                                steps).reshape(n, len(cd))
        self.Ts = field.transmittance_secondary(np.repeat(pts, len(sd), axis=0), np.tile(sd, (n, 1)), self.sig_e,
                                                steps).reshape(n, len(sd))

    def values(self, sun_pdf=None, albedo_power=1, shadow_env=False):
        """(E[L_sun]/sun colour, E[L_env]/env^2.2 for a constant env, E[alpha])."""
        p2 = sun_pdf_reference(self.cos_alpha) if sun_pdf is None else sun_pdf
        w = self.ph_c / (self.ph_c + p2)            # w2 for lrd; w1 for srd inside the cone
        a = self.alb ** albedo_power
        shadow = (self.Tc * w[None, :]) @ self.cw / self.omega
        g_sun = shadow + (self.Tc * (self.ph_c * w)[None, :]) @ self.cw
        g_env = (self.Ts * self.ph_s[None, :]) @ self.sw - (self.Tc * (self.ph_c * (1.0 - w))[None, :]) @ self.cw
        if shadow_env:
            g_env = g_env + shadow
        return a * float(self.prob @ g_sun), a * float(self.prob @ g_env), float(self.prob.sum())


def single_scatter(field, ro, rd, sigma_s, sigma_a, g, sun_dir, cos_alpha, **kw):
    return SingleScatter(field, ro, rd, sigma_s, sigma_a, g, sun_dir, cos_alpha, **kw).values()


def camera_dirs_pinhole(eye, target, fov_y_deg, w, h, px, py, up=(0.0, 0.0, 1.0)):
    """Primary ray directions of a pinhole camera from its parameters (no
    matrices): the reference's pixel-to-NDC map x_ndc = -(2(px+0.5)/w - 1),
    y_ndc = -(2(py+0.5)/h - 1) (kernel.cpp:462-463) applied to a camera
    looking from eye at target with +z up, vertical field of view fov_y."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    tn = np.tan(0.5 * np.radians(fov_y_deg))
    hx = -(2.0 * (np.asarray(px, np.float64) + 0.5) / w - 1.0)
    hy = -(2.0 * (np.asarray(py, np.float64) + 0.5) / h - 1.0)
    d = f[None, :] + (hx * tn * w / h)[:, None] * r[None, :] + (hy * tn)[:, None] * u[None, :]
    return d / np.linalg.norm(d, axis=1, keepdims=True)
