"""Seeded synthetic scenes for the solvers' inputs.

The reference takes its inputs only through ``LBMSolver3D.initialize(solid_np)``
and ``MPMSolver3D.initialize_particles_*`` (fsi/lbm3d.py:193-217,
fsi/mpm3d.py:69-165) and never uses ``SimulationConfig.random_seed``. These
generators build those inputs reproducibly from a seed
(``numpy.random.default_rng``): porous sphere packings for the lattice solid
mask, jittered particle lattices for boxes, spheres and ellipsoids.

All geometry is in lattice units with cell centres at integer + 0.5.
"""

from __future__ import annotations

import numpy as np


def porous_spheres(shape, porosity=0.8, rmin=4.0, rmax=10.0, seed=0, periodic=True):
    """Solid mask of random overlapping spheres, added until the solid fraction
    reaches 1 - porosity.

    Radii are U[rmin, rmax], centres U over the box; spheres wrap periodically
    when ``periodic``. A cell is solid when its centre lies inside a sphere.
    Returns (mask int32 (nx, ny, nz), info dict with the sphere list and the
    exact porosity).
    """
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(s) for s in shape)
    solid = np.zeros((nx, ny, nz), dtype=bool)
    target = (1.0 - porosity) * solid.size
    count = 0
    spheres = []
    while count < target:
        c = rng.uniform(0.0, 1.0, 3) * np.array([nx, ny, nz])
        r = rng.uniform(rmin, rmax)
        spheres.append((*c, r))
        lo = np.floor(c - r - 0.5).astype(int)
        hi = np.ceil(c + r - 0.5).astype(int) + 1
        axes = []
        idx = []
        for a, n in enumerate((nx, ny, nz)):
            cells = np.arange(lo[a], hi[a])
            if not periodic:
                cells = cells[(cells >= 0) & (cells < n)]
            axes.append((cells + 0.5 - c[a]) ** 2)
            idx.append(np.mod(cells, n))
        d2 = axes[0][:, None, None] + axes[1][None, :, None] + axes[2][None, None, :]
        inside = d2 <= r * r
        block = solid[np.ix_(*idx)]
        count += int(np.count_nonzero(inside & ~block))
        solid[np.ix_(*idx)] = block | inside
    info = {"spheres": np.array(spheres), "porosity": 1.0 - solid.mean(), "seed": seed}
    return solid.astype(np.int32), info


def channel_walls(shape, y=True, z=True):
    """Solid shell on the y and/or z faces (no-slip walls through the mask)."""
    solid = np.zeros(tuple(int(s) for s in shape), dtype=np.int32)
    if y:
        solid[:, 0, :] = 1
        solid[:, -1, :] = 1
    if z:
        solid[:, :, 0] = 1
        solid[:, :, -1] = 1
    return solid


def _lattice(lower, upper, spacing):
    axes = [np.arange(lower[a] + 0.5 * spacing, upper[a], spacing) for a in range(3)]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.stack([m.ravel() for m in mesh], axis=1)


def jittered_box(lower, upper, spacing, seed=0, jitter=0.5):
    """Particles on a lattice of the given spacing inside [lower, upper), each
    offset by U(-jitter/2, jitter/2) * spacing per axis."""
    rng = np.random.default_rng(seed)
    pos = _lattice(np.asarray(lower, float), np.asarray(upper, float), float(spacing))
    pos += rng.uniform(-0.5 * jitter, 0.5 * jitter, pos.shape) * spacing
    return pos.astype(np.float32)


def jittered_ellipsoid(center, semi_axes, spacing, rotation=None, seed=0, jitter=0.5):
    """Jittered lattice particles inside an ellipsoid (optionally rotated)."""
    center = np.asarray(center, float)
    semi = np.asarray(semi_axes, float)
    R = np.eye(3) if rotation is None else np.asarray(rotation, float)
    ext = np.abs(R) @ semi
    pos = jittered_box(center - ext, center + ext, spacing, seed=seed, jitter=jitter).astype(float)
    local = (pos - center) @ R
    keep = ((local / semi) ** 2).sum(axis=1) <= 1.0
    return pos[keep].astype(np.float32)


def jittered_sphere(center, radius, spacing, seed=0, jitter=0.5):
    return jittered_ellipsoid(center, (radius, radius, radius), spacing, seed=seed, jitter=jitter)


def random_rotation(rng):
    """Uniform random rotation matrix (QR of a Gaussian matrix with sign fix)."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1.0
    return q


def sphere_pack(shape, count, rmin, rmax, spacing=0.5, seed=0, margin=3.0, gap=1.0, jitter=0.5,
                max_tries=100_000):
    """Particles of `count` non-overlapping spheres in a box of `shape` cells.

    Radii are U[rmin, rmax]; centres U over the box shrunk by radius + margin
    on every face; a candidate is rejected when it comes closer than `gap`
    cells to an accepted sphere. Each sphere is a jittered lattice of the
    given spacing (seeded per sphere). Particles are returned sphere by
    sphere, each in lattice order.
    Returns (positions float32 (n, 3), info dict with centres, radii, counts).
    """
    rng = np.random.default_rng(seed)
    shape = np.asarray(shape, float)
    centres, radii = [], []
    tries = 0
    while len(centres) < count:
        tries += 1
        if tries > max_tries:
            raise ValueError(f"could not place {count} spheres in {tuple(shape)} after {max_tries} tries")
        r = rng.uniform(rmin, rmax)
        lo, hi = r + margin, shape - r - margin
        if np.any(hi <= lo):
            continue
        c = rng.uniform(lo, hi)
        if all(np.linalg.norm(c - c2) >= r + r2 + gap for c2, r2 in zip(centres, radii)):
            centres.append(c)
            radii.append(r)
    parts = [jittered_sphere(c, r, spacing, seed=seed * 1000 + i, jitter=jitter)
             for i, (c, r) in enumerate(zip(centres, radii))]
    pos = np.concatenate(parts, axis=0).astype(np.float32)
    info = {"centres": np.array(centres), "radii": np.array(radii), "counts": [len(p) for p in parts]}
    return pos, info


def body_moduli(counts, low, high, seed=0):
    """Young's modulus per body, log-uniform in [low, high] (numpy
    default_rng(seed)), repeated over each body's `counts` particles.
    Returns (per-particle float64 array, per-body array), for
    MPMSolver3D.set_particle_material with sphere_pack's info["counts"]."""
    rng = np.random.default_rng(seed)
    per_body = np.exp(rng.uniform(np.log(low), np.log(high), len(counts)))
    return np.repeat(per_body, np.asarray(counts, dtype=np.int64)), per_body
