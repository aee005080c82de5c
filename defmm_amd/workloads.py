"""Synthetic point clouds and charges of the benchmark configurations.

All randomness comes from SplitMix64 (Steele, Lea & Flood 2014) mapped to
53-bit doubles in [0, 1), so a (seed, size) pair names one exact input.

Configurations (wavenumber kappa = 2*pi*n_lambda / D with D = 2 * the
reference's box radius, i.e. twice the largest half-extent of the bounding
box, reference include/structs/particles.hpp:79-100):
  cube_edges  points on the surface of [-1/2, 1/2]^3, face chosen uniformly,
              face coordinates (u, v) by rejection with acceptance
              proportional to 1/(d_edge + 0.01), d_edge = min(u, 1-u, v, 1-v)
  sphere      uniform on the unit sphere (normalised Gaussian triples)
  ellipsoid   prolate (1, 1/4, 1/4) ellipsoid surface, Gaussian clusters
              (sigma 0.05) around seeded centres, radially projected
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n):
    """n successive SplitMix64 outputs (uint64) from the given seed."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n):
    """n doubles in [0, 1) (top 53 bits)."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n):
    """n standard normal doubles (Box-Muller on two uniform streams)."""
    u1 = uniform(seed, n)
    u2 = uniform(seed ^ 0x5DEECE66D, n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def sphere(n, seed=1, radius=1.0):
    g = np.stack([normal(seed + 11 * k, n) for k in range(3)], axis=1)
    return radius * g / np.linalg.norm(g, axis=1, keepdims=True)


def cube_edges(n, seed=1):
    """Surface of [-1/2, 1/2]^3, density ~ 1/(d_edge + 0.01) on each face."""
    out = np.empty((0, 3))
    s = seed
    amax = 1.0 / 0.01
    while len(out) < n:
        m = 4 * (n - len(out)) + 1024
        u = uniform(s, m)
        v = uniform(s + 1, m)
        a = uniform(s + 2, m)
        f = np.minimum((uniform(s + 3, m) * 6).astype(np.int64), 5)
        s += 4
        de = np.minimum(np.minimum(u, 1 - u), np.minimum(v, 1 - v))
        keep = a * amax < 1.0 / (de + 0.01)
        u, v, f = u[keep], v[keep], f[keep]
        axis = f // 2
        side = np.where(f % 2 == 0, 0.5, -0.5)
        p = np.empty((len(u), 3))
        for ax in range(3):
            sel = axis == ax
            o1, o2 = [k for k in range(3) if k != ax]
            p[sel, ax] = side[sel]
            p[sel, o1] = u[sel] - 0.5
            p[sel, o2] = v[sel] - 0.5
        out = np.concatenate([out, p])
    return out[:n]


def ellipsoid_clusters(n, seed=1, axes=(1.0, 0.25, 0.25), ncentres=8, sigma=0.05):
    """Prolate ellipsoid surface, Gaussian-clustered density."""
    a = np.asarray(axes, dtype=np.float64)
    c = sphere(ncentres, seed=seed + 101) * a
    k = np.minimum((uniform(seed + 7, n) * ncentres).astype(np.int64), ncentres - 1)
    g = np.stack([normal(seed + 13 * j + 3, n) for j in range(3)], axis=1)
    x = c[k] + sigma * g
    # radial projection onto the ellipsoid: x / |x / a|
    return x / np.linalg.norm(x / a, axis=1, keepdims=True)


def gaussian_clusters(n, seed=1, ncentres=16, sigma=0.05):
    """Gaussian clusters in the unit cube: centres uniform in [0, 1]^3, points
    centre + sigma * N(0, I3)."""
    c = np.stack([uniform(seed + 201 + k, ncentres) for k in range(3)], axis=1)
    k = np.minimum((uniform(seed + 9, n) * ncentres).astype(np.int64), ncentres - 1)
    g = np.stack([normal(seed + 17 * j + 5, n) for j in range(3)], axis=1)
    return c[k] + sigma * g


GEOMETRIES = {"sphere": sphere, "cube_edges": cube_edges, "ellipsoid": ellipsoid_clusters,
              "clusters": gaussian_clusters}


def charges(n, seed=7):
    """Complex charges with re, im ~ U(-1, 1)."""
    return (2 * uniform(seed, n) - 1) + 1j * (2 * uniform(seed + 1, n) - 1)


def box_diameter(*clouds):
    """2 * box radius of the reference (max half-extent of the joint bounding box)."""
    p = np.concatenate([np.asarray(c).reshape(-1, 3) for c in clouds])
    mn, mx = p.min(0), p.max(0)
    c = (mn + mx) / 2
    return 2.0 * max(np.max(c - mn), np.max(mx - c))


def kappa_for(n_lambda, *clouds):
    """Wavenumber putting n_lambda wavelengths across the box diameter."""
    return 2.0 * np.pi * n_lambda / box_diameter(*clouds)


CONFIGS = {
    # name: (geometry, n, n_lambda, L, Ncrit)  -- Ncrit = leaf size + 1
    "cube20k_32l": ("cube_edges", 20000, 32, 4, 65),
    # the same cube at order 8, the lowest order whose relative error against
    # the FP64 direct sum is <= 1e-6 there (7.2e-7; order 7 gives 5.8e-6)
    "cube20k_32l_L8": ("cube_edges", 20000, 32, 8, 65),
    "cube20k_32l_L9": ("cube_edges", 20000, 32, 9, 65),
    "cube20k_32l_L7": ("cube_edges", 20000, 32, 7, 65),
    "sphere1m_128l": ("sphere", 1000000, 128, 5, 65),
    # same sphere at 32 wavelengths with order 7: the relative-error <= 1e-6 regime
    "sphere1m_32l_L7": ("sphere", 1000000, 32, 7, 65),
    "sphere1m_32l_L8": ("sphere", 1000000, 32, 8, 65),
    "sphere1m_16l_L7": ("sphere", 1000000, 16, 7, 65),
    "ellipsoid1m_128l": ("ellipsoid", 1000000, 128, 5, 65),
    "ellipsoid2m_256l_L6": ("ellipsoid", 2000000, 256, 6, 97),
}


def make(name, seed=1):
    geo, n, nl, L, ncrit = CONFIGS[name]
    pts = GEOMETRIES[geo](n, seed)
    return pts, kappa_for(nl, pts), L, ncrit
