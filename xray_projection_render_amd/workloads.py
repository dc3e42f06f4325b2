"""Seeded synthetic workloads: the benchmark scene configurations.

Each builder returns plain data (scene dict in the reference's object
schema, optional deformation dict, view angles, resolution, ds) so the same
input drives libxrp, the FP64 oracle and the benchmark.

* ``csg``      - cube minus sphere minus cylinder, 256^2, 8 views (CPU-runnable case)
* ``voxel``    - 256^3 float32 Gaussian-blob grid on [-0.5,0.5]^3, 512^2, 64 views
* ``lattice``  - BCC cylinder lattice tessellated 8^3 with a linear
  deformation, 512^2, 180 views
* ``batch``    - 32 random primitive collections of uneven size
* ``gyroid``   - tessellated gyroid 6^3 with a sigmoid-profile z shear

Random numbers come from splitmix64 so every consumer sees the same values.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014); uniform() in [0, 1) from 53 bits."""

    def __init__(self, seed: int):
        self.state = seed & MASK

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def integer(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + self.next() % (hi - lo + 1)


@dataclass
class Workload:
    name: str
    obj: Optional[Dict]
    deformation: Optional[Dict]
    angles: List[Tuple[float, float]]
    resolution: int
    ds: float
    integrator: str = "simple"
    voxels: Optional[np.ndarray] = None   # (nz, nx, ny) float32
    R: float = 4.0
    fov: float = 40.0
    extra: Dict = field(default_factory=dict)

    @property
    def pixels(self) -> int:
        return len(self.angles) * self.resolution * self.resolution


def csg(resolution: int = 256, n_views: int = 8) -> Workload:
    """Unit cube (rho 1) minus a sphere r=0.3 and an x-cylinder r=0.15."""
    obj = {"type": "object_collection", "objects": [
        {"type": "cube", "center": [0.0, 0.0, 0.0], "side": 1.0, "rho": 1.0},
        {"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 0.3, "rho": -1.0},
        {"type": "cylinder", "p0": [-0.6, 0.0, 0.0], "p1": [0.6, 0.0, 0.0],
         "radius": 0.15, "rho": -1.0},
    ]}
    angles = [(360.0 * k / n_views, 90.0) for k in range(n_views)]
    # ds: the MinFeatureSize/5 rule of the reference (cube side 1 -> 0.1/5)
    return Workload("csg", obj, None, angles, resolution, 0.02)


def blob_grid(n: int = 256, n_blobs: int = 64, seed: int = 0) -> np.ndarray:
    """n^3 float32 densities, layout (z, x, y); node i sits at world
    -0.5 + i/(n-1).  rho = clip(sum_b a_b exp(-|x-c_b|^2 / (2 s_b^2)), 0, 1)."""
    rng = SplitMix64(seed)
    blobs = []
    for _ in range(n_blobs):
        c = [rng.uniform(-0.4, 0.4) for _ in range(3)]
        s = rng.uniform(0.02, 0.1)
        a = rng.uniform(0.2, 1.0)
        blobs.append((c, s, a))
    x = -0.5 + np.arange(n, dtype=np.float64) / (n - 1)
    acc = np.zeros((n, n, n), dtype=np.float64)  # (z, x, y)
    for (cx, cy, cz), s, a in blobs:
        k = 1.0 / (2.0 * s * s)
        ex = np.exp(-k * (x - cx) ** 2)
        ey = np.exp(-k * (x - cy) ** 2)
        ez = np.exp(-k * (x - cz) ** 2)
        acc += a * ez[:, None, None] * ex[None, :, None] * ey[None, None, :]
    return np.clip(acc, 0.0, 1.0).astype(np.float32)


VOXEL_DEFORMATION = {"type": "affine",
                     "matrix": [[2.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 2.0]]}


def voxel(resolution: int = 512, n_views: int = 64, n: int = 256,
          grid: Optional[np.ndarray] = None) -> Workload:
    """voxel_grid (fixed [-1,1]^3 support, objects.go:791) seen through the
    affine map diag(2,2,2), so the grid covers world [-0.5,0.5]^3."""
    rng = SplitMix64(1)
    angles = [(rng.uniform(0.0, 360.0), 90.0) for _ in range(n_views)]
    vox = blob_grid(n) if grid is None else grid
    ds = 2.0 / n / 5.0  # MinFeatureSize/5 for a voxel grid (objects.go:857-860)
    return Workload("voxel", None, dict(VOXEL_DEFORMATION), angles, resolution, ds,
                    voxels=vox)


def bcc_cell(radius: float = 0.06) -> Dict:
    corners = [[sx * 0.5, sy * 0.5, sz * 0.5] for sx in (-1, 1) for sy in (-1, 1)
               for sz in (-1, 1)]
    return {"type": "object_collection", "objects": [
        {"type": "cylinder", "p0": c, "p1": [0.0, 0.0, 0.0], "radius": radius, "rho": 1.0}
        for c in [[float(v) for v in cc] for cc in corners]]}


def lattice(resolution: int = 512, n_views: int = 180, cells: int = 8,
            ds: float = 0.0015) -> Workload:
    """BCC unit cell (8 corner-to-centre struts, r 0.06) tessellated over
    [-cells/2, cells/2]^3 and mapped into [-0.5,0.5]^3 by the affine
    deformation M = cells * (I + A), A_ij ~ U[-0.05, 0.05] (seed 2)."""
    rng = SplitMix64(2)
    A = [[rng.uniform(-0.05, 0.05) for _ in range(3)] for _ in range(3)]
    M = [[cells * ((1.0 if i == j else 0.0) + A[i][j]) for j in range(3)] for i in range(3)]
    h = cells / 2.0
    obj = {"type": "tessellated_obj_coll",
           "uc": {"objects": bcc_cell(), "xmin": -0.5, "xmax": 0.5, "ymin": -0.5,
                  "ymax": 0.5, "zmin": -0.5, "zmax": 0.5},
           "xmin": -h, "xmax": h, "ymin": -h, "ymax": h, "zmin": -h, "zmax": h}
    angles = [(2.0 * k * 180.0 / n_views, 90.0) for k in range(n_views)]
    return Workload("lattice", obj, {"type": "affine", "matrix": M}, angles, resolution, ds)


def _unit(rng: SplitMix64) -> List[float]:
    while True:
        v = [rng.uniform(-1.0, 1.0) for _ in range(3)]
        n = math.sqrt(sum(t * t for t in v))
        if 0.1 < n <= 1.0:
            return [t / n for t in v]


def random_scene(seed: int, max_prims: int = 200) -> Tuple[Dict, int, int]:
    """One scene of the batched workload; returns (object, resolution, views)."""
    rng = SplitMix64(3 + seed)
    n = rng.integer(1, max_prims)
    objs = []
    for _ in range(n):
        kind = ["sphere", "cylinder", "box", "parallelepiped", "gyroid"][rng.integer(0, 4)]
        c = [rng.uniform(-0.4, 0.4) for _ in range(3)]
        size = rng.uniform(0.02, 0.3)
        rho = rng.uniform(-1.0, 1.0)
        if kind == "sphere":
            objs.append({"type": "sphere", "center": c, "radius": size, "rho": rho})
        elif kind == "cylinder":
            u = _unit(rng)
            p0 = [c[k] - size * u[k] for k in range(3)]
            p1 = [c[k] + size * u[k] for k in range(3)]
            objs.append({"type": "cylinder", "p0": p0, "p1": p1, "radius": 0.5 * size,
                         "rho": rho})
        elif kind == "box":
            sides = [size, rng.uniform(0.02, 0.3), rng.uniform(0.02, 0.3)]
            objs.append({"type": "box", "center": c, "sides": sides, "rho": rho})
        elif kind == "parallelepiped":
            vs = []
            for _ in range(3):
                u = _unit(rng)
                ln = rng.uniform(0.02, 0.3)
                vs.append([ln * t for t in u])
            origin = [c[k] - 0.5 * (vs[0][k] + vs[1][k] + vs[2][k]) for k in range(3)]
            objs.append({"type": "parallelepiped", "origin": origin, "v0": vs[0],
                         "v1": vs[1], "v2": vs[2], "rho": rho})
        else:
            objs.append({"type": "gyroid", "center": c, "scale": rng.uniform(2.0, 8.0),
                         "thickness": rng.uniform(0.05, 0.2), "rho": rho})
    res = [128, 256, 512][rng.integer(0, 2)]
    views = rng.integer(4, 64)
    return {"type": "object_collection", "objects": objs}, res, views


def batch(n_scenes: int = 32, max_prims: int = 200) -> List[Workload]:
    out = []
    for s in range(n_scenes):
        obj, res, nv = random_scene(s, max_prims)
        rng = SplitMix64(1000 + s)
        angles = [(rng.uniform(0.0, 360.0), rng.uniform(5.0, 175.0)) for _ in range(nv)]
        # ds from the scene's own MinFeatureSize/5 rule is computed by the caller
        out.append(Workload(f"batch{s:02d}", obj, None, angles, res, -1.0))
    return out


TANH_SHEAR = {"type": "tanh_shear", "amplitude": 0.1, "rate": 8.0, "direction": "z",
              "source": "x"}


def gyroid_lattice(resolution: int = 1024, n_views: int = 256, cells: int = 6,
                   ds: float = 0.001, shear: bool = False) -> Workload:
    """Gyroid with one period per cell (scale = cell/(2 pi)), thickness 0.1,
    tessellated cells^3 over [-0.5,0.5]^3; with shear, the displacement
    z += 0.1 tanh(8 x) of the surrogate workload."""
    cell = 1.0 / cells
    obj = {"type": "tessellated_obj_coll",
           "uc": {"objects": {"type": "object_collection", "objects": [
               {"type": "gyroid", "center": [0.0, 0.0, 0.0],
                "scale": cell / (2.0 * math.pi), "thickness": 0.1, "rho": 1.0}]},
                  "xmin": -cell / 2, "xmax": cell / 2, "ymin": -cell / 2,
                  "ymax": cell / 2, "zmin": -cell / 2, "zmax": cell / 2},
           "xmin": -0.5, "xmax": 0.5, "ymin": -0.5, "ymax": 0.5, "zmin": -0.5, "zmax": 0.5}
    angles = [(360.0 * k / n_views, 90.0) for k in range(n_views)]
    return Workload("gyroid", obj, dict(TANH_SHEAR) if shear else None, angles, resolution,
                    ds)
