"""Seeded initial microstructures and the per-cell list format shared by every component.

A state is held in "CSR" form, the order of the raw PhaseField .dat files
(src/PhaseField.cpp:2340-2388): cells ordered x outer, z inner; `count[c]` entries per cell;
`index`/`value` the concatenated insertion-ordered lists.

`voronoi_grains` is this library's own generator (the reference's VoronoiTessellation,
src/Initializations.cpp:2014-2194, draws seeds from libstdc++'s implementation-defined
uniform_int_distribution, so its output is not reproduced; both sides of every comparison start
from the same file instead).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Lists:
    """Per-cell phase-field lists, CSR layout (see module docstring)."""

    shape: tuple  # storage shape (nx, ny, nz) with reduced axes = 1
    count: np.ndarray  # int32 [ncells]
    index: np.ndarray  # int32 [sum(count)]
    value: np.ndarray  # float64 [sum(count)]

    @property
    def ncells(self) -> int:
        return int(np.prod(self.shape))

    def offsets(self) -> np.ndarray:
        off = np.zeros(self.ncells + 1, dtype=np.int64)
        np.cumsum(self.count, out=off[1:])
        return off

    def to_slots(self, slots: int):
        """Slot-major arrays: ids[s, c], vals[s, c] (unused slots: id 0, value 0)."""
        n = self.ncells
        ids = np.zeros((slots, n), dtype=np.int32)
        vals = np.zeros((slots, n), dtype=np.float64)
        off = self.offsets()
        if self.count.max(initial=0) > slots:
            raise ValueError("list longer than the slot capacity")
        for s in range(int(self.count.max(initial=0))):
            cells = np.nonzero(self.count > s)[0]
            ids[s, cells] = self.index[off[cells] + s]
            vals[s, cells] = self.value[off[cells] + s]
        return ids, vals

    @staticmethod
    def from_slots(shape, count, ids, vals) -> "Lists":
        count = np.asarray(count, dtype=np.int32)
        n = count.shape[0]
        total = int(count.sum())
        off = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(count, out=off[1:])
        index = np.zeros(total, dtype=np.int32)
        value = np.zeros(total, dtype=np.float64)
        for s in range(int(count.max(initial=0))):
            cells = np.nonzero(count > s)[0]
            index[off[cells] + s] = ids[s, cells]
            value[off[cells] + s] = vals[s, cells]
        return Lists(tuple(shape), count, index, value)


def storage_shape(nx: int, ny: int, nz: int) -> tuple:
    """GridParameters::SetDimensions: a size of 0 marks a reduced axis of storage size 1."""
    return (nx or 1, ny or 1, nz or 1)


def voronoi_grains(nx: int, ny: int, nz: int, ngrains: int, seed: int, overlap: bool = True,
                   sides: str = "both") -> Lists:
    """Periodic Voronoi tessellation from `ngrains` uniform-random seed points.

    Each cell belongs to the grain of its nearest seed (periodic distance; ties resolved by the
    k-d tree). With `overlap`, every cell whose face neighbours (periodic) belong to other grains
    also lists those grains, in the order -x, +x, -y, +y, -z, +z, and all entries of the cell get
    the value 1/n: an overlap band that the first steps relax into the double-obstacle profile.
    `sides="plus"` looks at the +x, +y, +z neighbours only (a one-cell band: at most 1 + d
    entries per cell), `"both"` at all 2d face neighbours (a two-cell band).
    """
    from scipy.spatial import cKDTree

    shape = storage_shape(nx, ny, nz)
    active = [n > 0 for n in (nx, ny, nz)]
    rng = np.random.default_rng(seed)
    dims = [s for s, a in zip(shape, active) if a]
    seeds = rng.uniform(0.0, 1.0, size=(ngrains, len(dims))) * np.asarray(dims, dtype=np.float64)
    seeds = np.minimum(seeds, np.asarray(dims, dtype=np.float64) - 1e-9)
    tree = cKDTree(seeds, boxsize=np.asarray(dims, dtype=np.float64))
    axes = [np.arange(s, dtype=np.float64) for s, a in zip(shape, active) if a]
    grids = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([g.ravel() for g in grids], axis=1)
    grain = np.empty(pts.shape[0], dtype=np.int32)
    chunk = 1 << 21
    for lo in range(0, pts.shape[0], chunk):
        _, idx = tree.query(pts[lo:lo + chunk], k=1, workers=-1)
        grain[lo:lo + chunk] = idx
    g3 = grain.reshape(shape)
    if not overlap:
        n = g3.size
        return Lists(shape, np.ones(n, np.int32), g3.ravel().astype(np.int32), np.ones(n, np.float64))

    if sides not in ("both", "plus"):
        raise ValueError("sides must be 'both' or 'plus'")
    nbrs = []
    for ax in range(3):
        if not active[ax]:
            continue
        if sides == "both":
            nbrs.append(np.roll(g3, 1, axis=ax).ravel())   # neighbour at -1
        nbrs.append(np.roll(g3, -1, axis=ax).ravel())      # neighbour at +1
    own = g3.ravel()
    n = own.size
    cols = [own]
    for p, nb in enumerate(nbrs):
        keep = nb != own
        for q in range(p):
            keep &= nb != nbrs[q]
        cols.append(np.where(keep, nb, -1))
    table = np.stack(cols, axis=1)  # [n, 1 + 2d], -1 = no entry
    valid = table >= 0
    count = valid.sum(axis=1).astype(np.int32)
    index = table[valid].astype(np.int32)  # row-major: per cell in column order
    value = np.repeat(1.0 / count, count).astype(np.float64)
    return Lists(shape, count, index, value)


def random_rotations(ngrains: int, seed: int) -> np.ndarray:
    """Uniformly distributed orientations (Shoemake's method) as row-major rotation matrices,
    computed with the quaternion formula of Quaternion::setRotationMatrix
    (src/Containers/Quaternion.cpp:78-98). Returns (quaternions [n,4] (s,x,y,z), matrices [n,9])."""
    rng = np.random.default_rng(seed)
    u1, u2, u3 = rng.uniform(size=(3, ngrains))
    q = np.stack([
        np.sqrt(u1) * np.cos(2 * np.pi * u3),
        np.sqrt(1 - u1) * np.sin(2 * np.pi * u2),
        np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
        np.sqrt(u1) * np.sin(2 * np.pi * u3),
    ], axis=1)
    return q, quaternion_to_matrix(q)


def quaternion_to_matrix(q: np.ndarray) -> np.ndarray:
    q = np.asarray(q, dtype=np.float64)
    n = np.sqrt((q * q).sum(axis=1, keepdims=True))
    s, x, y, z = (q / n).T
    m = np.empty((q.shape[0], 9))
    m[:, 0] = 1.0 - 2.0 * (y * y + z * z)
    m[:, 1] = 2.0 * (x * y - s * z)
    m[:, 2] = 2.0 * (x * z + s * y)
    m[:, 3] = 2.0 * (x * y + s * z)
    m[:, 4] = 1.0 - 2.0 * (x * x + z * z)
    m[:, 5] = 2.0 * (y * z - s * x)
    m[:, 6] = 2.0 * (x * z - s * y)
    m[:, 7] = 2.0 * (y * z + s * x)
    m[:, 8] = 1.0 - 2.0 * (x * x + y * y)
    return m


@dataclass(frozen=True)
class CalibrationDomain:
    """One member of a calibration ensemble: a periodic 2D domain side x side with `ngrains`
    Voronoi grains (voronoi_grains(side, side, 0, ngrains, seed)) of one phase, isotropic
    interface energy `sigma` (J/m^2) and mobility `mu` (m^4/(J s))."""
    side: int
    ngrains: int
    sigma: float
    mu: float
    seed: int


def calibration_domains(n: int = 64, seed: int = 0x0FE6) -> list:
    """n domains with log-uniform sizes 128^2-512^2 (side round(128 * 4^U), U ~ U[0, 1]),
    log-uniform 20-300 grains, sigma ~ U[0.1, 0.5] and mu ~ logU[1e-9, 1e-8]; member i takes the
    Voronoi seed seed + i."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(round(128 * 4 ** rng.uniform()))
        ng = int(round(np.exp(rng.uniform(np.log(20), np.log(300)))))
        sig = float(rng.uniform(0.1, 0.5))
        mu = float(np.exp(rng.uniform(np.log(1e-9), np.log(1e-8))))
        out.append(CalibrationDomain(side, ng, sig, mu, seed + i))
    return out
