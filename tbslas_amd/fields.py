"""Analytic velocity and concentration fields (test fixtures and workload
inputs), vectorised with numpy.  pts is [n,3]; vector fields return [n,3],
scalars [n,1].

The reference's fields (src/utils/fields.h, applications/include/
field_wrappers.h) use PI = 3.14159265 (fields.h:16); these use the same
constant where they restate a reference field, and math.pi where a
workload definition states 2*pi.
"""
from __future__ import annotations

import math

import numpy as np

REF_PI = 3.14159265  # fields.h:16


def rotation(pts, omega: float = 1.0, xc: float = 0.5, yc: float = 0.5):
    """get_vorticity_field (fields.h:142-151): v = omega (yc - y, x - xc, 0)."""
    p = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    v = np.zeros_like(p)
    v[:, 0] = omega * (yc - p[:, 1])
    v[:, 1] = omega * (p[:, 0] - xc)
    return v


def rotation_axis(pts, axis, omega: float, centre=(0.5, 0.5, 0.5)):
    """Rigid rotation about a unit axis through centre: v = omega a x (p - c)."""
    p = np.asarray(pts, dtype=np.float64).reshape(-1, 3) - np.asarray(centre)
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    return omega * np.cross(a, p)


def gaussian(pts, centre=(0.6, 0.5, 0.5), sigma=0.06, amp: float = 1.0):
    """get_gaussian_field_3d (fields.h:275-293) with equal sigmas."""
    p = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    s = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (3,))
    c = np.asarray(centre, dtype=np.float64)
    e = ((p[:, 0] - c[0]) ** 2) / (2 * s[0] * s[0]) + ((p[:, 1] - c[1]) ** 2) / (
        2 * s[1] * s[1]) + ((p[:, 2] - c[2]) ** 2) / (2 * s[2] * s[2])
    return (amp * np.exp(-e))[:, None]


def gaussian_at(pts, t: float, omega: float = 1.0, centre0=(0.6, 0.5, 0.5),
                sigma: float = 0.06, axis_centre=(0.5, 0.5)):
    """get_gaussian_field_atT (field_wrappers.h:31-56): the blob carried by
    the rigid rotation omega about (0.5, 0.5) for time t."""
    xc0, yc0, zc0 = centre0
    theta0 = math.atan2(yc0 - axis_centre[1], xc0 - axis_centre[0])
    r = math.hypot(xc0 - axis_centre[0], yc0 - axis_centre[1])
    th = theta0 + omega * t
    c = (axis_centre[0] + r * math.cos(th), axis_centre[1] + r * math.sin(th), zc0)
    return gaussian(pts, c, sigma)


def slotted_sphere(pts, centre=(0.5, 0.5, 0.5), R: float = 0.3, w: float = 0.1,
                   a: float = 0.0):
    """get_slotted_cylinder (fields.h:220-242): a slotted sphere indicator."""
    p = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    xc, yc, zc = centre
    r2 = (p[:, 0] - xc) ** 2 + (p[:, 1] - yc) ** 2 + (p[:, 2] - zc) ** 2
    out = (r2 < R * R).astype(np.float64)
    d = math.cos(a) * (p[:, 1] - yc) - math.sin(a) * (p[:, 0] - xc)
    x = math.sin(a) * (p[:, 1] - yc) + math.cos(a) * (p[:, 0] - xc)
    out[(r2 < R * R) & (np.abs(d) < w) & (x + w > 0)] = 0.0
    return out[:, None]


def taylor_green_2d3(pts):
    """Spatial part of the time-dependent Taylor-Green field of the adaptive
    workload: (sin 2pi x cos 2pi y cos 2pi z, -cos 2pi x sin 2pi y cos 2pi z, 0);
    the full field is this times cos(pi t)."""
    p = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    tp = 2.0 * math.pi
    sx, cx = np.sin(tp * p[:, 0]), np.cos(tp * p[:, 0])
    sy, cy = np.sin(tp * p[:, 1]), np.cos(tp * p[:, 1])
    cz = np.cos(tp * p[:, 2])
    v = np.zeros_like(p)
    v[:, 0] = sx * cy * cz
    v[:, 1] = -cx * sy * cz
    return v


def taylor_green_time_factor(t: float) -> float:
    return math.cos(math.pi * t)


def taylor_green_tv(pts, t: float, freq: float = 2 * REF_PI, A=1.0, B=1.0, C=1.0,
                    a=2 * REF_PI, b=2 * REF_PI, c=2 * REF_PI):
    """get_taylor_green_field_tv (fields.h:84-113)."""
    p = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    f = math.cos(freq * t)
    v = np.empty_like(p)
    v[:, 0] = A * np.cos(a * p[:, 0]) * np.sin(b * p[:, 1]) * np.sin(c * p[:, 2]) * f
    v[:, 1] = B * np.sin(a * p[:, 0]) * np.cos(b * p[:, 1]) * np.sin(c * p[:, 2]) * f
    v[:, 2] = C * np.sin(a * p[:, 0]) * np.sin(b * p[:, 1]) * np.cos(c * p[:, 2]) * f
    return v


def lamb_oseen_pair(pts, centres=((0.4, 0.5), (0.6, 0.5)), core: float = 0.05,
                    gamma: float = 1.0):
    """Two co-rotating Lamb-Oseen vortex tubes along z: each contributes the
    azimuthal speed gamma / (2 pi r) (1 - exp(-r^2 / core^2))."""
    p = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    v = np.zeros_like(p)
    for cx, cy in centres:
        dx, dy = p[:, 0] - cx, p[:, 1] - cy
        r2 = dx * dx + dy * dy
        with np.errstate(divide="ignore", invalid="ignore"):
            f = np.where(r2 > 0, gamma / (2 * math.pi * r2) * (1 - np.exp(-r2 / (core * core))),
                         0.0)
        v[:, 0] += -dy * f
        v[:, 1] += dx * f
    return v


def smoothed_slotted_sphere(pts, width: float, centre=(0.5, 0.5, 0.5), R: float = 0.3,
                            w: float = 0.1):
    """Slotted sphere indicator smoothed with tanh profiles of half-width
    `width` across the sphere surface and the slot walls."""
    p = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    xc, yc, zc = centre
    r = np.sqrt((p[:, 0] - xc) ** 2 + (p[:, 1] - yc) ** 2 + (p[:, 2] - zc) ** 2)
    ball = 0.5 * (1 - np.tanh((r - R) / width))
    dy = np.abs(p[:, 1] - yc)
    in_slot_y = 0.5 * (1 - np.tanh((dy - w) / width))
    in_slot_x = 0.5 * (1 + np.tanh((p[:, 0] - xc + w) / width))
    return (ball * (1 - in_slot_y * in_slot_x))[:, None]


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n uniform doubles in [0, 1) from the SplitMix64 sequence of `seed`
    (u = (z >> 11) * 2^-53): the same numbers on every platform."""
    mask = (1 << 64) - 1
    out = np.empty(n)
    s = seed & mask
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * 2.0 ** -53
    return out


def random_poly_velocity(q: int, seed: int = 4):
    """A smooth random velocity on [0,1]^3: per component one global
    Chebyshev polynomial of total degree q, coefficients c_m = u_m /
    (1 + |m|)^3 with u_m uniform in [-1, 1) from SplitMix64(seed), |m| =
    i + j + k, rescaled by 1 / sum |c_m| so that |v_k| <= 1.  Restricted to
    the leaves of any tree of degree >= q it is exact, so it is continuous
    across leaves.  Returns f(pts [n,3]) -> [n,3]."""
    tri = [(i, j, k) for i in range(q + 1) for j in range(q + 1 - i) for k in range(q + 1 - i - j)]
    nc = len(tri)
    u = 2.0 * splitmix64(seed, 3 * nc) - 1.0
    deg = np.array([i + j + k for i, j, k in tri], dtype=np.float64)
    c = u.reshape(3, nc) / (1.0 + deg) ** 3
    c = c / np.abs(c).sum(axis=1, keepdims=True)

    def f(pts):
        p = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
        xi = 2.0 * p - 1.0
        T = [np.polynomial.chebyshev.chebvander(xi[:, a], q) for a in range(3)]  # [n, q+1]
        out = np.zeros((len(p), 3))
        for m, (i, j, k) in enumerate(tri):
            b = T[2][:, i] * T[1][:, j] * T[0][:, k]
            out += b[:, None] * c[:, m][None, :]
        return out

    return f


def tracer_blobs(n: int = 64, seed: int = 3):
    """n Gaussian tracer blobs: centres uniform in [0.3, 0.7]^3 and sigma
    uniform in [0.03, 0.08], drawn from SplitMix64(seed) in the order
    (cx, cy, cz, sigma) per tracer.  Returns (centres [n,3], sigma [n])."""
    u = splitmix64(seed, 4 * n).reshape(n, 4)
    return 0.3 + 0.4 * u[:, :3], 0.03 + 0.05 * u[:, 3]


def gaussian_blobs(pts, centres, sigma):
    """Values [m, n] of n Gaussian blobs exp(-|x - c|^2 / (2 sigma^2)) at pts
    [m, 3]; works on numpy arrays and torch tensors alike."""
    d2 = ((pts[:, None, :] - centres[None, :, :]) ** 2).sum(-1)
    return (-d2 / (2 * sigma[None, :] ** 2)).exp() if hasattr(d2, "exp") else np.exp(
        -d2 / (2 * sigma[None, :] ** 2))


def rotation_coefficients(anchor, depth, q: int, axis, omega: float,
                          centre=(0.5, 0.5, 0.5)) -> np.ndarray:
    """Chebyshev coefficients [L, 3, Nc] of the rigid rotation v = omega a x
    (p - c) on every leaf, in closed form: v is linear, so on a leaf with
    corner k and edge h, with p = k + h (xi + 1) / 2, only T_0 (the value at
    the leaf centre) and the three T_1 terms (omega a x e_d h / 2 for axis
    d) are nonzero.  Packed order i(z) -> j(y) -> k(x): T_1(x) at index 1,
    T_1(y) at index q + 1, T_1(z) at index (q + 1)(q + 2) / 2."""
    anchor = np.asarray(anchor, dtype=np.float64).reshape(-1, 3)
    h = 0.5 ** np.asarray(depth, dtype=np.float64)
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    d = q + 1
    nc = d * (d + 1) * (d + 2) // 6
    out = np.zeros((len(h), 3, nc))
    mid = anchor * h[:, None] + 0.5 * h[:, None] - np.asarray(centre)
    out[:, :, 0] = omega * np.cross(a, mid)
    for ax, idx in ((0, 1), (1, d), (2, d * (d + 1) // 2)):
        e = np.zeros(3)
        e[ax] = 1.0
        out[:, :, idx] = omega * np.cross(a, e)[None, :] * (0.5 * h)[:, None]
    return out
