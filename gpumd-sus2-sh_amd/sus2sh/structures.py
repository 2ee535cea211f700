"""Synthetic atomic structures for the benchmark and parity configurations.

Positions are returned as float64 arrays of shape [3, N] (GPUMD SoA order x[N], y[N], z[N])
and cells as 3x3 arrays whose ROWS are the cell vectors a, b, c.
"""
import numpy as np


def make_box(cell, pbc=(1, 1, 1)):
    """GPUMD box array h[18]: h[0..8] = matrix with the cell vectors as columns (row-major),
    h[9..17] = its inverse (src/force/force.cu:439-465 uses the same convention)."""
    cell = np.asarray(cell, dtype=np.float64)
    H = cell.T.copy()
    Hi = np.linalg.inv(H)
    h = np.concatenate([H.reshape(-1), Hi.reshape(-1)])
    return h, tuple(int(p) for p in pbc)


def fcc(nx, ny, nz, a):
    """Cubic fcc supercell: 4 nx ny nz atoms, cell diag(nx a, ny a, nz a)."""
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    pos = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    cell = np.diag([nx * a, ny * a, nz * a])
    return pos.T.copy(), cell


def assign_species(n, fractions, seed):
    """Exact species counts by fraction (remainder to the last species), shuffled by seed."""
    counts = [int(round(f * n)) for f in fractions[:-1]]
    counts.append(n - sum(counts))
    types = np.concatenate([np.full(c, t, dtype=np.int32) for t, c in enumerate(counts)])
    rng = np.random.default_rng(seed)
    rng.shuffle(types)
    return types


def jitter(pos, sigma, seed):
    rng = np.random.default_rng(seed)
    return pos + rng.normal(0.0, sigma, size=pos.shape)


def fcc_alloy(nx, ny, nz, a, fractions, sigma, species_seed, jitter_seed):
    pos, cell = fcc(nx, ny, nz, a)
    n = pos.shape[1]
    types = assign_species(n, fractions, species_seed)
    pos = jitter(pos, sigma, jitter_seed)
    return pos, types, cell


def random_min_separation(n, density, min_sep, fractions, seed, max_rounds=200):
    """Random sequential addition in a periodic cube of side (n/density)^(1/3), rejecting
    candidates closer than min_sep to any accepted atom (minimum image)."""
    rng = np.random.default_rng(seed)
    side = (n / density) ** (1.0 / 3.0)
    ncell = max(1, int(side // min_sep))
    width = side / ncell
    grid = {}
    accepted = np.zeros((n, 3))
    count = 0
    rounds = 0
    while count < n and rounds < max_rounds:
        rounds += 1
        cand = rng.uniform(0.0, side, size=(max(1024, 2 * (n - count)), 3))
        for p in cand:
            c = tuple((p // width).astype(int) % ncell)
            ok = True
            for dx in (-1, 0, 1):
                if not ok:
                    break
                for dy in (-1, 0, 1):
                    if not ok:
                        break
                    for dz in (-1, 0, 1):
                        key = ((c[0] + dx) % ncell, (c[1] + dy) % ncell, (c[2] + dz) % ncell)
                        for q in grid.get(key, ()):
                            d = accepted[q] - p
                            d -= side * np.round(d / side)
                            if d @ d < min_sep * min_sep:
                                ok = False
                                break
                        if not ok:
                            break
            if ok:
                accepted[count] = p
                grid.setdefault(c, []).append(count)
                count += 1
                if count == n:
                    break
    if count < n:
        raise RuntimeError("random_min_separation: could not place all atoms")
    types = assign_species(n, fractions, seed + 1)
    return accepted.T.copy(), types, np.diag([side, side, side])


def triclinic_cell(n, density, rng, angle_range=(75.0, 105.0)):
    """Random triclinic cell of volume n/density with angles alpha, beta, gamma drawn
    uniformly from angle_range (degrees) and edge lengths scaled isotropically."""
    for _ in range(100):
        al, be, ga = np.radians(rng.uniform(*angle_range, size=3))
        # lower-triangular cell from lengths 1 and angles
        a = np.array([1.0, 0.0, 0.0])
        b = np.array([np.cos(ga), np.sin(ga), 0.0])
        cx = np.cos(be)
        cy = (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
        cz2 = 1.0 - cx * cx - cy * cy
        if cz2 <= 1e-3:
            continue
        c = np.array([cx, cy, np.sqrt(cz2)])
        cell = np.stack([a, b, c])
        vol = abs(np.linalg.det(cell))
        scale = (n / density / vol) ** (1.0 / 3.0)
        return cell * scale
    raise RuntimeError("triclinic_cell: no valid angle triple")
