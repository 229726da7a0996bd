"""Seeded synthetic configurations of the benchmark workloads.

Fibers: orientations uniform on the sphere (normalised Gaussian), centres
uniform in a ball (c = R U^(1/3) h/|h|), rejected when any node lies within
`min_sep` of a node of an already placed fiber (node-node separation).  Nodes
are c + (s_m - L/2) t with s_m = (1 + alpha_m) L/2 on the discretisation's
parameter grid.  The periphery is the Fibonacci sphere of
shape_gallery.py:7-32 with inward normals (tstep.py:293) and equal quadrature
weights 4 pi R^2 / N.
"""
from dataclasses import dataclass, field

import numpy as np
from scipy.spatial import cKDTree

from .geometry import shape_gallery_ellipsoid, shape_gallery_sphere


@dataclass
class FiberSpec:
    num_fibers: int
    num_points: object          # int, or a list of allowed node counts
    length: float
    ball_radius: float
    E: float = 0.01
    epsilon: float = 1e-3
    min_sep: float = 0.2
    end_force: float = 1.0       # plus-end force along the end tangent


@dataclass
class Config:
    name: str
    seed: int
    fibers: FiberSpec
    periphery_nodes: int = 0
    periphery_radius: float = 0.0
    eta: float = 1.0
    dt: float = 1e-3
    penalty: float = 500.0
    nfd: int = 4
    discretisation: str = "reference"   # uniform nodes + FD ('reference') or 'chebyshev'
    gmres_tol: float = 1e-8
    restart: int = 150
    maxiter: int = 200
    steps: int = 1
    extra: dict = field(default_factory=dict)


def config(i, discretisation="reference"):
    """The five workloads (0: small FP64 plain, 2: batched LU, 3: time series, 4: large)."""
    if i == 0:
        return Config("small", 0, FiberSpec(32, 32, 1.0, 4.0), 1024, 5.0,
                      discretisation=discretisation)
    if i == 2:
        return Config("batched_lu", 2, FiberSpec(2048, [16, 24, 32, 48, 64], 1.0, 20.0),
                      discretisation=discretisation)
    if i == 3:
        return Config("time_series", 3, FiberSpec(1000, 32, 1.0, 8.0), 8000, 10.0, steps=400,
                      discretisation=discretisation)
    if i == 4:
        return Config("large", 4, FiberSpec(4096, 64, 1.5, 14.0), 32768, 16.0, steps=50,
                      discretisation=discretisation)
    raise ValueError(f"no fiber-system config {i}")


def _alpha(N, discretisation):
    if discretisation == "reference":
        return np.linspace(-1.0, 1.0, N)
    return -np.cos(np.pi * np.arange(N) / (N - 1))


class _NodeIndex:
    """Nearest-node queries over the nodes placed so far: a k-d tree of the
    bulk (rebuilt every 16384 nodes), one of the recent nodes (rebuilt every
    512) and a brute-force check of the newest few hundred."""

    def __init__(self):
        self.bulk = np.zeros((0, 3))
        self.recent = np.zeros((0, 3))
        self.pending = []
        self.t_bulk = None
        self.t_recent = None

    def too_close(self, x, sep):
        for t in (self.t_bulk, self.t_recent):
            if t is not None and t.query(x, k=1, distance_upper_bound=sep)[0].min() < sep:
                return True
        if self.pending:
            P = np.concatenate(self.pending)
            if np.linalg.norm(x[:, None, :] - P[None, :, :], axis=2).min() < sep:
                return True
        return False

    def add(self, x):
        self.pending.append(x)
        if sum(p.shape[0] for p in self.pending) > 512:
            self.recent = np.concatenate([self.recent] + self.pending)
            self.pending = []
            if self.recent.shape[0] > 16384:
                self.bulk = np.concatenate([self.bulk, self.recent])
                self.recent = np.zeros((0, 3))
                self.t_bulk = cKDTree(self.bulk)
                self.t_recent = None
            else:
                self.t_recent = cKDTree(self.recent)


def make_fibers(spec: FiberSpec, seed: int, discretisation="reference"):
    """List of (N_k, 3) node arrays placed without node-node overlap."""
    rng = np.random.default_rng(seed)
    placed = []
    index = _NodeIndex()
    attempts = 0
    while len(placed) < spec.num_fibers:
        attempts += 1
        if attempts > 200 * spec.num_fibers:
            raise RuntimeError("fiber placement failed: packing too dense")
        if isinstance(spec.num_points, (list, tuple)):
            N = int(rng.choice(spec.num_points))
        else:
            N = int(spec.num_points)
        g = rng.normal(size=3)
        t = g / np.linalg.norm(g)
        h = rng.normal(size=3)
        c = spec.ball_radius * rng.uniform() ** (1.0 / 3.0) * h / np.linalg.norm(h)
        s = (1.0 + _alpha(N, discretisation)) * spec.length / 2.0
        x = c[None, :] + (s - spec.length / 2.0)[:, None] * t[None, :]
        if index.too_close(x, spec.min_sep):
            continue
        placed.append(x)
        index.add(x)
    return placed


def make_periphery(n, radius):
    """Shell nodes, inward normals, equal weights."""
    x, nout = shape_gallery_sphere(n, radius)
    w = np.full(n, 4.0 * np.pi * radius * radius / n)
    return x, -nout, w


def make_periphery_ellipsoid(n, a, b, c):
    """Ellipsoidal shell of semi-axes (a, b, c): nodes and normals of the
    reference's ellipsoid gallery, normals turned inward, and weights that
    carry each Fibonacci cell's equal area 4 pi / N on the unit sphere onto
    the ellipsoid: w = (4 pi / N) a b c |diag(1/a, 1/b, 1/c) u| for the
    node's pre-image u (the area factor of the linear map).  The reference
    computes its weights with its RBF quadrature module instead."""
    x, nout = shape_gallery_ellipsoid(n, a, b, c)
    u = x / np.array([a, b, c])
    w = (4.0 * np.pi / n) * a * b * c * np.sqrt((u[:, 0] / a) ** 2 + (u[:, 1] / b) ** 2
                                                + (u[:, 2] / c) ** 2)
    return x, -nout, w


def make_sphere_body(centre, radius=0.5, nodes=600):
    """Rigid sphere: Fibonacci nodes (shape_gallery.py:7-32), outward normals,
    equal quadrature weights 4 pi a^2 / N."""
    x, n = shape_gallery_sphere(nodes, radius)
    return (np.asarray(centre, float), x + np.asarray(centre, float)[None, :], n,
            np.full(nodes, 4.0 * np.pi * radius * radius / nodes))


def make_bodies(count, radius, nodes, ball_radius, fibers=(), margin=0.1, seed=0):
    """Spheres with centres uniform in a ball, no overlap with each other or with
    fiber nodes (clearance `margin`)."""
    rng = np.random.default_rng(seed)
    fib = np.concatenate(fibers) if len(fibers) else np.zeros((0, 3))
    tree = cKDTree(fib) if len(fib) else None
    centres = []
    for _ in range(10000 * max(count, 1)):
        if len(centres) == count:
            break
        h = rng.normal(size=3)
        c = ball_radius * rng.uniform() ** (1.0 / 3.0) * h / np.linalg.norm(h)
        if any(np.linalg.norm(c - d) < 2 * radius + margin for d in centres):
            continue
        if tree is not None and tree.query(c)[0] < radius + margin:
            continue
        centres.append(c)
    if len(centres) < count:
        raise RuntimeError("body placement failed")
    return [make_sphere_body(c, radius, nodes) for c in centres]
