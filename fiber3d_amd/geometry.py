"""Host-side geometry generators (one-off setup, NumPy).

shape_gallery_sphere restates shape_gallery.shape_gallery('sphere', ...)
(python/shape_gallery/shape_gallery.py:7-32): a Fibonacci lattice on the
sphere with outward normals.
"""
import numpy as np


def shape_gallery_sphere(number_of_nodes, radius):
    """Fibonacci-lattice nodes (N,3) and outward unit normals (N,3).

    Node i + N/2 (i in [-N/2, N/2)) sits at latitude asin(2i/(2*(N//2)+1)) and
    longitude (i mod phi) * 2 pi / phi wrapped into [-pi, pi], as in
    shape_gallery.py:12-24.
    """
    phi = (1.0 + np.sqrt(5.0)) / 2.0
    half = number_of_nodes // 2
    nodes = np.zeros((number_of_nodes, 3))
    i = np.arange(-half, half)
    lat = np.arcsin((2.0 * i) / (2 * half + 1))
    lon = np.mod(i, phi) * 2.0 * np.pi / phi
    lon = np.where(lon < -np.pi, 2.0 * np.pi + lon, lon)
    lon = np.where(lon > np.pi, lon - 2.0 * np.pi, lon)
    nodes[i + half, 0] = np.cos(lon) * np.cos(lat)
    nodes[i + half, 1] = np.sin(lon) * np.cos(lat)
    nodes[i + half, 2] = np.sin(lat)
    nodes *= radius
    normals = 2.0 * nodes
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    return nodes, normals


def shape_gallery_ellipsoid(number_of_nodes, a, b, c):
    """Nodes (N,3) and outward unit normals (N,3) of the ellipsoid
    x^2/a^2 + y^2/b^2 + z^2/c^2 = 1: the Fibonacci lattice of
    shape_gallery_sphere with its coordinates scaled by (a, b, c), normals
    along the level-set gradient 2 p / (a^2, b^2, c^2)
    (shape_gallery.py:34-68)."""
    u, _ = shape_gallery_sphere(number_of_nodes, 1.0)
    abc = np.array([a, b, c], dtype=np.float64)
    nodes = u * abc
    normals = 2.0 * nodes / abc ** 2
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    return nodes, normals
