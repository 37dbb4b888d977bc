"""Seeded synthetic particle sets for the DC-PSE workloads.

Every generator is deterministic (numpy default_rng with a fixed seed) and
returns host arrays, so the GPU path and the oracle see identical inputs.

  lattice_2d_periodic   2D 64x64 lattice, x=(i+1/2)h, h=1/64, rCut=3.1h, periodic
  jittered_2d           2D n x n lattice + U(-0.1h, 0.1h) jitter, rCut=3.1h, periodic
                        (512^2: ~29 neighbours; 1024^2 for the adjoint batch)
  jittered_3d_holes     3D 128^3 lattice + U(-0.1h, 0.1h) jitter, 10% removed at
                        random, rCut=2.4h, periodic (~1.89M particles, 35..60 neighbours)
  adaptive_3d           3D particles with spacing h(x) = h0 (1 + 3x) along x,
                        per-particle rCut = 2.4 h(x_p)
  uneven_3d             the same construction with h(x) = h0 (1 + x) and one
                        rCut for all particles, so supports grow from about 25
                        neighbours at the coarse end to about 200 at the fine
                        end (batched uneven-size assembly)
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2.0 * np.pi


def _lattice(n, dim):
    h = 1.0 / n
    g = (np.arange(n) + 0.5) * h
    axes = np.meshgrid(*([g] * dim), indexing="ij")
    return np.stack([a.ravel() for a in axes], axis=1), h


def lattice_2d_periodic(n=64):
    pos, h = _lattice(n, 2)
    return dict(pos=pos, h=h, rcut=3.1 * h, box_lo=[0.0, 0.0], box_hi=[1.0, 1.0], periodic=0b11)


def jittered_2d(n=512, seed=2, jitter=0.1, rcut_h=3.1):
    pos, h = _lattice(n, 2)
    rng = np.random.default_rng(seed)
    pos = pos + rng.uniform(-jitter * h, jitter * h, size=pos.shape)
    pos = np.mod(pos, 1.0)
    return dict(pos=pos, h=h, rcut=rcut_h * h, box_lo=[0.0, 0.0], box_hi=[1.0, 1.0], periodic=0b11)


def jittered_3d_holes(n=128, seed=3, jitter=0.1, drop=0.1, rcut_h=2.4):
    pos, h = _lattice(n, 3)
    rng = np.random.default_rng(seed)
    pos = pos + rng.uniform(-jitter * h, jitter * h, size=pos.shape)
    keep = rng.random(pos.shape[0]) >= drop
    pos = np.mod(pos[keep], 1.0)
    return dict(pos=np.ascontiguousarray(pos), h=h, rcut=rcut_h * h, box_lo=[0.0] * 3, box_hi=[1.0] * 3,
                periodic=0b111)


def adaptive_3d(n_target=1_000_000, seed=5, h_ratio=4.0, rcut_h=2.4, jitter=0.1):
    """Spacing grows linearly from h0 at x=0 to h_ratio*h0 at x=1 (periodic in y, z).

    x-planes are placed at x_{k+1} = x_k + h(x_k); each plane is an m x m
    lattice in (y, z) with m = round(1/h(x_k)), so the local spacing is
    isotropic. h0 is chosen so the total count is close to n_target.
    """
    rng = np.random.default_rng(seed)

    def planes(h0):
        xs, hs = [], []
        x = 0.5 * h0
        while x < 1.0:
            hx = h0 * (1.0 + (h_ratio - 1.0) * x)
            xs.append(x)
            hs.append(hx)
            x += hx
        return np.array(xs), np.array(hs)

    lo, hi = 1e-3, 0.1
    for _ in range(60):
        h0 = np.sqrt(lo * hi)
        xs, hs = planes(h0)
        cnt = sum(int(round(1.0 / hh)) ** 2 for hh in hs)
        if cnt > n_target:
            lo = h0
        else:
            hi = h0
    xs, hs = planes(h0)
    chunks, hp = [], []
    for x, hh in zip(xs, hs):
        m = max(int(round(1.0 / hh)), 4)
        g = (np.arange(m) + 0.5) / m
        Y, Z = np.meshgrid(g, g, indexing="ij")
        P = np.stack([np.full(m * m, x), Y.ravel(), Z.ravel()], axis=1)
        P += rng.uniform(-jitter * hh, jitter * hh, size=P.shape)
        chunks.append(P)
        hp.append(np.full(m * m, hh))
    pos = np.concatenate(chunks)
    pos[:, 1:] = np.mod(pos[:, 1:], 1.0)
    hvec = np.concatenate(hp)
    # x is not periodic: the box is padded by the largest spacing on both sides
    return dict(pos=np.ascontiguousarray(pos), h=hvec, rcut=rcut_h * hvec,
                box_lo=[-h_ratio * h0, 0.0, 0.0], box_hi=[1.0 + h_ratio * h0, 1.0, 1.0], periodic=0b110)


def sin_field(pos):
    """f = prod_d sin(2 pi x_d) and its analytic d/dx and Laplacian."""
    s = np.sin(TWO_PI * pos)
    c = np.cos(TWO_PI * pos)
    f = np.prod(s, axis=1)
    dx = TWO_PI * c[:, 0] * np.prod(s[:, 1:], axis=1)
    dy = TWO_PI * c[:, 1] * s[:, 0] * (np.prod(s[:, 2:], axis=1) if pos.shape[1] > 2 else 1.0)
    lap = -pos.shape[1] * TWO_PI ** 2 * f
    return f, dx, dy, lap


def uneven_3d(n_target=1_000_000, seed=6, h_ratio=2.1, n_min=25, jitter=0.1):
    """adaptive_3d's particles (spacing from h0 to h_ratio * h0 along x) under a
    single cut-off sized so the coarsest particles keep about n_min
    neighbours: rcut = (3 n_min / (4 pi))^(1/3) * h_max, which gives the finest
    about h_ratio^3 * n_min (25 -> about 200 at h_ratio 2.1)."""
    w = adaptive_3d(n_target, seed=seed, h_ratio=h_ratio, jitter=jitter)
    h_max = float(np.max(w["h"]))
    rcut = (3.0 * n_min / (4.0 * np.pi)) ** (1.0 / 3.0) * h_max
    # periodic in x too (the fine end meets the coarse one across the wrap), so
    # no particle has a one-sided support with fewer than n_min neighbours
    pos = w["pos"].copy()
    pos[:, 0] = np.mod(pos[:, 0], 1.0)
    return dict(pos=pos, h=w["h"], rcut=rcut, box_lo=[0.0, 0.0, 0.0], box_hi=[1.0, 1.0, 1.0], periodic=0b111)

