"""Synthetic scenes and environments of the benchmark configurations.

Rows are N x 14 float32: centre(3), forward quaternion w,x,y,z (4),
scale(3), albedo(3), mass(1) -- the vp_primitive layout of include/vp_api.h.
Mass is the PLY convention of the reference (sigma_t_0 x sigma multiplier is
the total mass M, src/thesis/host/utils/io/ply.cpp:107-116); a peak density
rho converts as M = rho (2 pi)^(3/2) prod(s) (test/scenes/geometric_validation.cpp:17-19).

Every generator is seeded and deterministic.
"""
import math

import numpy as np

TWO_PI_POW_1_5 = (2.0 * math.pi) ** 1.5
SUN_DIR = np.array([0.7197, 0.4530, 0.5261])  # test/scenes/single_gaussian.cpp:86-88
ERF_3_OVER_SQRT2 = math.erf(3.0 / math.sqrt(2.0))


def random_quaternions(rng, n):
    """Uniform rotations (Shoemake 1992, 'Uniform random rotations')."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    x = a * np.sin(2 * np.pi * u2)
    y = a * np.cos(2 * np.pi * u2)
    z = b * np.sin(2 * np.pi * u3)
    w = b * np.cos(2 * np.pi * u3)
    return np.stack([w, x, y, z], axis=1)


def rows(centers, quats, scales, albedo, mass):
    n = centers.shape[0]
    out = np.empty((n, 14), np.float32)
    out[:, 0:3] = centers
    out[:, 3:7] = quats
    out[:, 7:10] = scales
    out[:, 10:13] = albedo
    out[:, 13] = mass
    return out


def c0_single_gaussian(albedo=0.0, scale=0.25, central_tau=2.0):
    """One isotropic Gaussian at the origin whose central-ray optical depth,
    integrated over the 3-sigma support, equals central_tau:
    M = central_tau * 2 pi s^2 / erf(3 / sqrt 2)."""
    mass = central_tau * 2.0 * math.pi * scale * scale / ERF_3_OVER_SQRT2
    return rows(np.zeros((1, 3)), np.array([[1.0, 0, 0, 0]]), np.full((1, 3), scale),
                np.full((1, 3), albedo), np.array([mass]))


def c1_scattering(n=4096, seed=1, sigma_multiplier=10.0, albedo=0.9):
    """Centres U[-1,1]^3, per-axis scale log-U[0.02, 0.1], uniform rotations,
    peak density lognormal(0, 0.5) times the sigma multiplier."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-1.0, 1.0, (n, 3))
    scales = np.exp(rng.uniform(math.log(0.02), math.log(0.1), (n, 3)))
    quats = random_quaternions(rng, n)
    rho = np.exp(rng.normal(0.0, 0.5, n))
    mass = rho * TWO_PI_POW_1_5 * np.prod(scales, axis=1) * sigma_multiplier
    return rows(centers, quats, scales, np.full((n, 3), albedo), mass)


def _value_noise(p, seed):
    """Lattice value noise with quintic smoothstep, trilinear (values in [0,1])."""
    i = np.floor(p).astype(np.int64)
    f = p - i
    u = f * f * f * (f * (f * 6.0 - 15.0) + 10.0)

    def h(ix, iy, iz):
        x = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        x = (x ^ (x >> 13)) * 1274126177
        x = x ^ (x >> 16)
        return (x & 0xFFFFFF).astype(np.float64) / float(0xFFFFFF)

    c = {}
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                c[(dx, dy, dz)] = h(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz)
    x00 = c[(0, 0, 0)] * (1 - u[:, 0]) + c[(1, 0, 0)] * u[:, 0]
    x10 = c[(0, 1, 0)] * (1 - u[:, 0]) + c[(1, 1, 0)] * u[:, 0]
    x01 = c[(0, 0, 1)] * (1 - u[:, 0]) + c[(1, 0, 1)] * u[:, 0]
    x11 = c[(0, 1, 1)] * (1 - u[:, 0]) + c[(1, 1, 1)] * u[:, 0]
    y0 = x00 * (1 - u[:, 1]) + x10 * u[:, 1]
    y1 = x01 * (1 - u[:, 1]) + x11 * u[:, 1]
    return y0 * (1 - u[:, 2]) + y1 * u[:, 2]


def fbm(p, seed, octaves=5, base=2.0, gain=0.5, lacunarity=2.0):
    """5-octave fBm of value noise normalised to [0, 1]."""
    total, amp, freq, norm = 0.0, 1.0, base, 0.0
    for o in range(octaves):
        total = total + amp * _value_noise(p * freq, seed + o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def c2_cloud(n=1_000_000, seed=2, sigma_multiplier=7.5, albedo=0.99):
    """Rejection-sample centres in [-1,1]x[-0.5,0.5]^2 with acceptance fBm(x);
    per-axis scale log-U[0.005, 0.03]; density = fBm(x) times the multiplier."""
    rng = np.random.default_rng(seed)
    pts, dens = [], []
    have = 0
    while have < n:
        cand = rng.uniform([-1.0, -0.5, -0.5], [1.0, 0.5, 0.5], (max(2 * (n - have), 1024), 3))
        f = fbm(cand, seed)
        keep = rng.random(cand.shape[0]) < f
        pts.append(cand[keep])
        dens.append(f[keep])
        have += int(keep.sum())
    centers = np.concatenate(pts)[:n]
    rho = np.concatenate(dens)[:n]
    scales = np.exp(rng.uniform(math.log(0.005), math.log(0.03), (n, 3)))
    quats = random_quaternions(rng, n)
    mass = rho * TWO_PI_POW_1_5 * np.prod(scales, axis=1) * sigma_multiplier
    return rows(centers, quats, scales, np.full((n, 3), albedo), mass)


def torus_shell(n=250_000, seed=3, R=1.0, r=0.3, sigma_multiplier=10.0, albedo=0.95):
    """Centres uniform on the torus surface (area-correct rejection on the
    tube angle), per-axis scale log-U[0.01, 0.04], lognormal peak density."""
    rng = np.random.default_rng(seed)
    out = []
    have = 0
    while have < n:
        m = 2 * (n - have) + 1024
        th = rng.uniform(0, 2 * np.pi, m)
        ph = rng.uniform(0, 2 * np.pi, m)
        keep = rng.random(m) < (R + r * np.cos(ph)) / (R + r)
        th, ph = th[keep], ph[keep]
        out.append(np.stack([(R + r * np.cos(ph)) * np.cos(th), r * np.sin(ph), (R + r * np.cos(ph)) * np.sin(th)], 1))
        have += int(keep.sum())
    centers = np.concatenate(out)[:n]
    scales = np.exp(rng.uniform(math.log(0.01), math.log(0.04), (n, 3)))
    quats = random_quaternions(rng, n)
    rho = np.exp(rng.normal(0.0, 0.5, n))
    mass = rho * TWO_PI_POW_1_5 * np.prod(scales, axis=1) * sigma_multiplier
    return rows(centers, quats, scales, np.full((n, 3), albedo), mass)


def random_scene(n, seed, extent=1.0, smin=0.05, smax=0.3, albedo=0.8, rho_scale=3.0):
    """Small random scenes for parity tests."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-extent, extent, (n, 3))
    scales = np.exp(rng.uniform(math.log(smin), math.log(smax), (n, 3)))
    quats = random_quaternions(rng, n)
    rho = rho_scale * np.exp(rng.normal(0.0, 0.5, n))
    mass = rho * TWO_PI_POW_1_5 * np.prod(scales, axis=1)
    alb = rng.uniform(0.2, 1.0, (n, 3)) * albedo
    return rows(centers, quats, scales, alb, mass)


def constant_env(value=(1.0, 1.0, 1.0), w=64, h=32):
    env = np.empty((h, w, 4), np.float32)
    env[..., 0:3] = value
    env[..., 3] = 1.0
    return env


def texel_dirs(w, h, su=0.5, sv=0.5):
    """Direction of texel (u, v) at sub-texel offset (su, sv), the mapping of
    device/core/sampling.cuh:197-205."""
    u = (np.arange(w) + su) / w
    v = (np.arange(h) + sv) / h
    theta = 2.0 * np.pi * u[None, :] - np.pi
    polar = np.pi * v[:, None]
    return np.stack([np.sin(polar) * np.cos(theta), np.cos(polar) * np.ones_like(theta),
                     np.sin(polar) * np.sin(theta)], axis=-1)


def sky_env(w=2048, h=1024, sun_radius_deg=0.25, sun_scale=1e4, supersample=16):
    """Sky gradient over polar angle (zenith (0.25,0.45,1) -> horizon (1,0.95,0.9),
    ground 0.3 x (0.5,0.45,0.4)) plus a sun disk of the given angular radius
    along the meadow sun direction with radiance sun_scale x zenith; sun texel
    coverage supersampled."""
    zen = np.array([0.25, 0.45, 1.0])
    hor = np.array([1.0, 0.95, 0.9])
    gnd = 0.3 * np.array([0.5, 0.45, 0.4])
    d = texel_dirs(w, h)
    y = d[..., 1:2]
    t = np.sqrt(np.clip(y, 0.0, 1.0))
    sky = hor * (1.0 - t) + zen * t
    col = np.where(y > 0.0, sky, gnd)
    sun = SUN_DIR / np.linalg.norm(SUN_DIR)
    cos_r = math.cos(math.radians(sun_radius_deg))
    near = np.argwhere((d @ sun) > math.cos(math.radians(sun_radius_deg) + 2.5 * math.pi / h))
    for (vv, uu) in near:
        offs = (np.arange(supersample) + 0.5) / supersample
        su, sv = np.meshgrid(offs, offs)
        theta = 2.0 * np.pi * (uu + su) / w - np.pi
        polar = np.pi * (vv + sv) / h
        dd = np.stack([np.sin(polar) * np.cos(theta), np.cos(polar), np.sin(polar) * np.sin(theta)], -1)
        cover = float(np.mean((dd @ sun) >= cos_r))
        col[vv, uu] = col[vv, uu] * (1.0 - cover) + sun_scale * zen * cover
    env = np.empty((h, w, 4), np.float32)
    env[..., 0:3] = col
    env[..., 3] = 1.0
    return env


def default_camera_args(width, height, dist=3.5, fov=40.0, view="negz"):
    """asset_validation camera (test/scenes/cloud_validation.cpp:114-127)."""
    eye = {"negz": (0, 0, -dist), "posz": (0, 0, dist), "posx": (dist, 0, 0), "negx": (-dist, 0, 0),
           "diag": (dist * 0.6, dist * 0.5, -dist * 0.6)}[view]
    return dict(w=width, h=height, eye=np.array(eye, np.float32), target=np.zeros(3, np.float32),
                up=np.array([0, 1, 0], np.float32), vfov=fov)


# Per-scene render settings of the job mix: (phase g, generator)
MIX_SCENES = ("scattering", "cloud", "torus")
MIX_PHASE_G = {"scattering": 0.0, "cloud": 0.85, "torus": 0.0}


def mix_jobs(n_jobs=64, seed=4, res_div=1, spp_div=1):
    """The batched job mix: resolution uniform over {64,...,1024}^2, spp =
    round(exp U[ln 4, ln 256]), scene uniform over (scattering 4096,
    cloud 1M, torus 250k), camera on a sphere of radius 3.5 (uniform
    direction) looking at the origin with a 40 degree vertical FOV.
    res_div / spp_div shrink every job by the same factor (spp >= 1).
    Returns dicts: scene, width, height, spp, seed, phase_g, camera args."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n_jobs):
        res = int(rng.choice([64, 128, 256, 512, 1024]))
        spp = int(round(np.exp(rng.uniform(np.log(4.0), np.log(256.0)))))
        sc = MIX_SCENES[int(rng.integers(0, 3))]
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        up = np.array([0, 1, 0], np.float32) if abs(v[1]) < 0.99 else np.array([1, 0, 0], np.float32)
        w = max(8, res // res_div)
        jobs.append(dict(scene=sc, width=w, height=w, spp=max(1, spp // spp_div), seed=i,
                         phase_g=MIX_PHASE_G[sc],
                         cam=dict(w=w, h=w, eye=(3.5 * v).astype(np.float32), target=np.zeros(3, np.float32),
                                  up=up, vfov=40.0)))
    return jobs
