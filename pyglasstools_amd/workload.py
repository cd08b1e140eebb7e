"""Seeded synthetic glass configurations for the IK stress-field workloads.

Recipe (all draws from numpy Generator(PCG64(seed)), seed = 1000*config + frame):
  box       square/cubic, L = (N/rho)^(1/d), rho = 1; 2D boxes have z-size 1,
            non-periodic in z (as pyglasstools/SimBox.cc:12-18 sets them up)
  lattice   n_side = ceil(N^(1/d)) sites per axis at spacing L/n_side, offset
            by half a spacing; a seeded random subset of N sites; Gaussian
            jitter of 0.1 per coordinate; wrapped into [-L/2, L/2)
  diameters P(sigma) ~ sigma^-3 on [0.73, 1.62] by inverse CDF:
            sigma = (a - u (a - b))^(-1/2), a = 0.73^-2, b = 1.62^-2
  mass      1 (what gsd.hoomd reports when a file has no mass chunk)
  velocity  Maxwell-Boltzmann N(0, T/m) per component at T = 0.1, centre-of-
            mass velocity removed, v_z = 0 in 2D
  grid      numpy.linspace(-L/2, L/2, n) per axis, x outermost (the ordering
            of pyglasstools/irvingkirkwood/__init__.py:46-51)
"""
import math

import numpy as np

SIGMA_MIN, SIGMA_MAX = 0.73, 1.62


def make_system(N, dim, seed, rho=1.0, jitter=0.1, T=0.1):
    rng = np.random.Generator(np.random.PCG64(seed))
    L = (N / rho) ** (1.0 / dim)
    n_side = int(math.ceil(N ** (1.0 / dim) - 1e-9))
    while n_side ** dim < N:
        n_side += 1
    spacing = L / n_side
    nsites = n_side ** dim
    sites = rng.permutation(nsites)[:N] if nsites > N else np.arange(N)
    idx = np.stack(np.unravel_index(sites, (n_side,) * dim), axis=1).astype(np.float64)
    pos = np.zeros((N, 3))
    pos[:, :dim] = -0.5 * L + (idx + 0.5) * spacing
    pos[:, :dim] += rng.normal(0.0, jitter, size=(N, dim))
    pos[:, :dim] -= L * np.floor((pos[:, :dim] + 0.5 * L) / L)
    # keep strictly inside [-L/2, L/2) after rounding
    pos[:, :dim] = np.where(pos[:, :dim] >= 0.5 * L, pos[:, :dim] - L, pos[:, :dim])
    a, b = SIGMA_MIN ** -2, SIGMA_MAX ** -2
    u = rng.random(N)
    diam = (a - u * (a - b)) ** -0.5
    mass = np.ones(N)
    vel = np.zeros((N, 3))
    vel[:, :dim] = rng.normal(0.0, math.sqrt(T), size=(N, dim)) / np.sqrt(mass)[:, None]
    vel[:, :dim] -= vel[:, :dim].mean(axis=0)
    if dim == 2:
        Lvec = np.array([L, L, 1.0])
        periodic = np.array([1, 1, 0])
    else:
        Lvec = np.array([L, L, L])
        periodic = np.array([1, 1, 1])
    return {"dim": dim, "L": Lvec, "periodic": periodic, "pos": pos, "diam": diam,
            "mass": mass, "vel": vel, "seed": seed}


def cartesian_grid_2d(L, n):
    p = np.linspace(-L / 2.0, L / 2.0, n)
    g = np.zeros((n * n, 3))
    g[:, 0] = np.repeat(p, n)
    g[:, 1] = np.tile(p, n)
    return g


def cartesian_grid_3d(L, n):
    p = np.linspace(-L / 2.0, L / 2.0, n)
    X, Y, Z = np.meshgrid(p, p, p, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


# The workloads named by the benchmark configurations.
CONFIGS = {
    1: {"N": 2000, "dim": 2, "frames": 1, "grid": 32},
    2: {"N": 65536, "dim": 2, "frames": 100, "grid": 256},
    3: {"N": 131072, "dim": 3, "frames": 20, "grid": 64},
    5: {"N": [4096, 8192, 16384, 24576, 32768, 49152, 65536, 98304], "dim": 2, "frames": 1000, "grid": 128},
}


def config_frame(config, frame, N=None):
    c = CONFIGS[config]
    n = c["N"] if N is None else N
    sysd = make_system(n, c["dim"], 1000 * config + frame)
    L = sysd["L"][0]
    grid = cartesian_grid_2d(L, c["grid"]) if c["dim"] == 2 else cartesian_grid_3d(L, c["grid"])
    return sysd, grid
