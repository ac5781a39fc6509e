"""Seeded synthetic inputs for the projection workloads (host-side setup).

The 3-D two-phase dam break at n^3 on the unit domain (h = 1/n):
  * liquid block x < 0.4, y < 0.6 seeded with 2x2x2 jittered particles per
    cell (computeFlipBoxSeed3, grid_flip_solver3.js:70-130).  The jitter
    (<= 0.2 sub-spacing) keeps every particle inside its own cell, so the
    occupied cells are exactly the block's round(0.4 n) x round(0.6 n) x n
    cells: that is what the liquid fraction below uses (the GPU binning of the
    seeded particles reproduces it, see the transfer tests);
  * two-phase density rho = rho_gas + (rho_liq - rho_gas) * fraction, face
    beta = rho_ref / max(rho_face, 1e-4) with the arithmetic face mean
    (grid_two_phase_flip_solver2.js:693-762, in 3-D), ratio 1000:1;
  * the first frame's velocity u* = P2G(particles at rest) + g dt on every
    face, walls closed (zero normal velocity on the domain faces);
  * the singular all-Neumann system pinned at the top-centre cell
    (pressurePin 'topCenter', grid_two_phase_flip_solver2.js:133-169, 770-796).
All float arithmetic is f32, in the order of the reference expressions.
"""
import math

import numpy as np


def js_round(v):
    return int(np.floor(v + 0.5))


def liquid_block_cells(n, box_max=(0.4, 0.6, 1.0)):
    h = 1.0 / n
    return tuple(max(1, js_round(b / h)) for b in box_max)


def dam_break_fraction(n, box_max=(0.4, 0.6, 1.0)):
    cx, cy, cz = liquid_block_cells(n, box_max)
    frac = np.zeros((n, n, n), np.float32)  # [k][j][i]
    frac[:cz, :cy, :cx] = 1.0
    return frac.ravel()


def particle_count(n, ppa=2, box_max=(0.4, 0.6, 1.0)):
    cx, cy, cz = liquid_block_cells(n, box_max)
    return cx * cy * cz * ppa ** 3


def two_phase_beta(frac, shape, rho_gas=1.0, rho_liq=1000.0):
    """face beta per axis, flat first-index-fastest arrays"""
    nx, ny, nz = shape
    f = frac.reshape(nz, ny, nx)
    min_d = np.float32(1e-4)
    r0 = max(np.float32(rho_gas), min_d)
    r1 = max(np.float32(rho_liq), min_d)
    ref = max(r0, r1)
    rho = (r0 + (r1 - r0) * f).astype(np.float32)
    out = []
    for a in range(3):
        ax = 2 - a  # numpy axis of grid axis a
        m = rho.shape[ax]
        idx_hi = np.clip(np.arange(m + 1), 0, m - 1)
        idx_lo = np.clip(np.arange(m + 1) - 1, 0, m - 1)
        ra = np.take(rho, idx_lo, axis=ax)
        rb = np.take(rho, idx_hi, axis=ax)
        rf = ((ra + rb) * np.float32(0.5)).astype(np.float32)
        out.append((ref / np.maximum(rf, min_d)).astype(np.float32).ravel())
    return out


def dam_break_system(n, dt=1.0 / 60.0, g=-9.8, rho_gas=1.0, rho_liq=1000.0, box_max=(0.4, 0.6, 1.0)):
    shape = (n, n, n)
    frac = dam_break_fraction(n, box_max)
    beta = two_phase_beta(frac, shape, rho_gas, rho_liq)
    U = np.zeros((n, n, n + 1), np.float32)
    V = np.full((n, n + 1, n), np.float32(g) * np.float32(dt), np.float32)
    W = np.zeros((n + 1, n, n), np.float32)
    V[:, 0, :] = 0.0
    V[:, n, :] = 0.0
    active = np.zeros((n, n, n), np.uint8)
    active[n // 2, n - 1, n // 2] = 1
    return {
        "shape": shape,
        "spacing": (1.0 / n,) * 3,
        "fraction": frac,
        "beta": beta,
        "vel": [U.ravel(), V.ravel(), W.ravel()],
        "active": active.ravel(),
        "target": np.zeros(n ** 3, np.float32),
        "particles": particle_count(n, 2, box_max),
    }


def random_scene(n, seed):
    """A 128^3-style randomized variable-coefficient scene: random spheres and
    boxes of denser fluid, band-limited random divergence, closed walls."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    frac = np.zeros((n, n, n), np.float32)
    for _ in range(rng.integers(1, 7)):
        c = rng.uniform(0.2, 0.8, 3) * n
        r = rng.uniform(6, 30) * n / 128.0
        if rng.random() < 0.5:
            inside = (i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2 < r * r
        else:
            inside = (np.abs(i - c[0]) < r) & (np.abs(j - c[1]) < r) & (np.abs(k - c[2]) < r)
        frac[inside] = 1.0
    ratio = float(np.exp(rng.uniform(0, np.log(1000.0))))
    beta = two_phase_beta(frac.ravel(), shape, 1.0, ratio)
    vel = [rng.standard_normal(int(np.prod(s))).astype(np.float32) for s in
           ((n + 1) * n * n, n * (n + 1) * n, n * n * (n + 1))]
    for a, v in enumerate(vel):
        fs = [n, n, n]
        fs[a] += 1
        arr = v.reshape(fs[::-1])
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[2 - a] = 0
        sl1[2 - a] = -1
        arr[tuple(sl0)] = 0
        arr[tuple(sl1)] = 0
    active = np.zeros(n ** 3, np.uint8)
    active[(n // 2) + n * ((n - 1) + n * (n // 2))] = 1
    return {"shape": shape, "spacing": (1.0 / n,) * 3, "fraction": frac.ravel(), "beta": beta, "vel": vel,
            "active": active, "target": np.zeros(n ** 3, np.float32), "ratio": ratio}


def mixed_precision_scene(n, seed, band=0.25):
    """A scene of the mixed-precision configuration: a random union of 1-6
    spheres / boxes (radius 6-30 cells at 128^3) of the denser phase, the
    density ratio log-uniform in [1, 1000], a band-limited Gaussian-noise
    velocity (spectrum cut at `band` of the Nyquist wavenumber) whose
    divergence is the RHS, closed walls and an open top layer of cells
    (Dirichlet 0, which keeps the system well posed)."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    frac = np.zeros((n, n, n), np.float32)
    for _ in range(rng.integers(1, 7)):
        c = rng.uniform(0.2, 0.8, 3) * n
        r = rng.uniform(6, 30) * n / 128.0
        if rng.random() < 0.5:
            inside = (i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2 < r * r
        else:
            inside = (np.abs(i - c[0]) < r) & (np.abs(j - c[1]) < r) & (np.abs(k - c[2]) < r)
        frac[inside] = 1.0
    ratio = float(np.exp(rng.uniform(0, np.log(1000.0))))
    beta = two_phase_beta(frac.ravel(), shape, 1.0, ratio)
    vel = []
    for a in range(3):
        fs = [n, n, n]
        fs[a] += 1
        noise = rng.standard_normal(fs[::-1])
        spec = np.fft.rfftn(noise)
        kz = np.abs(np.fft.fftfreq(fs[2]))[:, None, None]
        ky = np.abs(np.fft.fftfreq(fs[1]))[None, :, None]
        kx = np.abs(np.fft.rfftfreq(fs[0]))[None, None, :]
        spec[(kx > 0.5 * band) | (ky > 0.5 * band) | (kz > 0.5 * band)] = 0
        arr = np.fft.irfftn(spec, s=fs[::-1], axes=(0, 1, 2)).astype(np.float32)
        arr /= max(float(np.abs(arr).max()), 1e-30)
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[2 - a] = 0
        sl1[2 - a] = -1
        arr[tuple(sl0)] = 0
        arr[tuple(sl1)] = 0
        vel.append(arr.ravel())
    active = (j.ravel() == n - 1).astype(np.uint8)
    return {"shape": shape, "spacing": (1.0 / n,) * 3, "fraction": frac.ravel(), "beta": beta, "vel": vel,
            "active": active, "target": np.zeros(n ** 3, np.float32), "ratio": ratio}


def mulberry32(seed, count):
    """createJitterRandom (grid/jitter_random.js:29-40): `count` doubles in [0, 1)"""
    out = np.empty(count, np.float64)
    state = seed & 0xFFFFFFFF
    m = 0xFFFFFFFF
    for q in range(count):
        state = (state + 0x6D2B79F5) & m
        t = state
        t = ((t ^ (t >> 15)) * (t | 1)) & m
        t ^= (t + (((t ^ (t >> 7)) * (t | 61)) & m)) & m
        out[q] = ((t ^ (t >> 14)) & m) / 4294967296.0
    return out


def circle_polygon(cx, cy, radius, segments=48):
    """makeCirclePolygon (examples/16-karman-vortex-street/main.js:217-229)"""
    return [(cx + radius * math.cos((i / segments) * math.pi * 2), cy + radius * math.sin((i / segments) * math.pi * 2))
            for i in range(segments)]


def wall_strip_polygon(inner_x, outer_x, ny, margin=1000.0):
    """makeWallStripPolygon (examples/16-karman-vortex-street/main.js:236-240)"""
    return [(inner_x, -margin), (outer_x, -margin), (outer_x, ny + margin), (inner_x, ny + margin)]


def karman2d(nx=256, ny=128, radius=12.0, center=(64.0, 64.0), segments=48, inflow_speed=1.0, noise=1e-3,
             seed=0x9E3779B9, frame_dt=1.0, courant=0.5, levels=4, tolerance=1e-6):
    """The 2-D Karman vortex street, plain method: a 48-gon cylinder collider,
    an inflow strip (x < 2, u set to inflow_speed), an outflow slab
    (x > nx - 2, Dirichlet p = 0 and the convective velocity condition),
    closed top and bottom (free slip), u = inflow_speed + U(-noise, noise)
    per U face from mulberry32(seed) in index order, v = 0; h = 1; MGPCG
    with `levels` levels and 2/2 red-black sweeps to `tolerance`; each frame
    of frame_dt splits into sub-steps of Courant number <= courant
    (scene layout after examples/16-karman-vortex-street/main.js:189-400)."""
    r = mulberry32(seed, (nx + 1) * ny)
    u0 = (inflow_speed + (r * 2.0 - 1.0) * noise).astype(np.float32)
    return {
        "resolution": (nx, ny), "spacing": (1.0, 1.0), "origin": (0.0, 0.0),
        "collider": [circle_polygon(center[0], center[1], radius, segments)],
        "inflow": [wall_strip_polygon(2.0, -1000.0, ny)], "inflow_velocity": (inflow_speed, 0.0),
        "outflow": [wall_strip_polygon(nx - 2.0, nx + 1000.0, ny)],
        "closed_flags": 4 | 8, "advect_collider": True,
        "u0": u0, "v0": np.zeros(nx * (ny + 1), np.float32),
        "frame_dt": frame_dt, "courant": courant, "max_sub_steps": 32,
        "multigrid": {"levels": levels, "nu_down": 2, "nu_up": 2, "n_coarsest": 20}, "tolerance": tolerance,
    }


def flip_dam_break(n, ppa=2, box_max=(0.4, 0.6, 1.0), jitter=0.2, seed=0):
    """Particles of the two-phase FLIP dam break on the unit domain at n^3
    (h = 1/n): ppa^3 particles per cell in EVERY cell (the gas is simulated
    too, as in grid_two_phase_flip_solver2.js), on the sub-cell lattice of
    computeFlipBoxSeed3 (grid_flip_solver3.js:70-130) with a jitter of
    +-jitter sub-spacings per axis (numpy's generator, not mulberry32: the
    full-size scene only needs a deterministic seed); concentration 1 in the
    liquid block (cells with centre x < 0.4, y < 0.6) and 0 in the gas.
    Returns float32 positions (x, y, z per particle) and concentrations,
    generated one sub-plane at a time."""
    h = 1.0 / n
    m = n * ppa
    sub = h / ppa
    cx, cy, _ = liquid_block_cells(n, box_max)
    rng = np.random.default_rng(seed)
    count = m ** 3
    pos = np.empty((count, 3), np.float32)
    conc = np.empty(count, np.float32)
    ii, jj = np.meshgrid(np.arange(m, dtype=np.float32), np.arange(m, dtype=np.float32), indexing="xy")
    ii, jj = ii.ravel(), jj.ravel()
    liquid_xy = ((ii // ppa) < cx) & ((jj // ppa) < cy)
    per = m * m
    for kk in range(m):
        sl = slice(kk * per, (kk + 1) * per)
        jit = rng.uniform(-jitter, jitter, (per, 3)).astype(np.float32)
        pos[sl, 0] = (ii + np.float32(0.5) + jit[:, 0]) * np.float32(sub)
        pos[sl, 1] = (jj + np.float32(0.5) + jit[:, 1]) * np.float32(sub)
        pos[sl, 2] = (np.float32(kk) + np.float32(0.5) + jit[:, 2]) * np.float32(sub)
        conc[sl] = liquid_xy.astype(np.float32)
    return {"n": (n, n, n), "spacing": (h, h, h), "positions": pos.ravel(), "concentration": conc, "count": count,
            "liquid": int(conc.sum()), "ratio": 1000.0}


def smoke_plume_sources(n, radius, center=None, noise=0.01, seed=0):
    """The 3-D smoke plume's per-cell sources on an n^3 grid of unit cells:
    a ball of `radius` cells (cell centres (i + 0.5, j + 0.5, k + 0.5),
    strictly inside, the example's test in f32) at the bottom centre
    (n/2, radius + 1.5, n/2) by default, density 1 and temperature
    1 + U(-noise, noise) per cell from `seed`; 0 outside."""
    c = np.array(center if center is not None else (n / 2.0, radius + 1.5, n / 2.0), np.float32)
    x = np.arange(n, dtype=np.float32) + np.float32(0.5)
    dx = (x - c[0])[None, None, :]
    dy = (x - c[1])[None, :, None]
    dz = (x - c[2])[:, None, None]
    inside = ((dx * dx + dy * dy) + dz * dz) < np.float32(radius) * np.float32(radius)
    rng = np.random.default_rng(seed)
    jitter = rng.uniform(-noise, noise, (n, n, n)).astype(np.float32)
    dens = np.where(inside, np.float32(1.0), np.float32(0.0)).astype(np.float32)
    temp = np.where(inside, np.float32(1.0) + jitter, np.float32(0.0)).astype(np.float32)
    return dens.ravel(), temp.ravel()
