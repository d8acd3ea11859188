"""Periodic connectivity of binary cells, for the binary check of a fiber design (nlfiber/topology.py:8-59):
connected solid components on the torus, the spanning test, and the removal of floating pieces.

Components are the connected components of a graph whose vertices are the solid voxels and whose edges join
periodic neighbours (26-neighbourhood in 3D, 8 in 2D, unless a connectivity stencil is given): the edges come
from shifting the cell by every half-stencil offset, and scipy's sparse connected_components labels the graph.
Labels are numbered 1..n in the order of each component's first voxel in C order (0 is the background), which
is the numbering the reference's relabelling produces.
"""
import itertools

import numpy as np
from scipy import ndimage
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components


def _offsets(d, connectivity):
    """Half of the neighbourhood stencil (one of each +/- pair), as integer offsets."""
    st = np.ones((3,) * d, dtype=bool) if connectivity is None else np.asarray(connectivity, dtype=bool)
    if st.shape != (3,) * d:
        raise ValueError(f"connectivity must be a {(3,) * d} boolean stencil")
    half = []
    for off in itertools.product((-1, 0, 1), repeat=d):
        if not any(off) or not st[tuple(o + 1 for o in off)]:
            continue
        first = next(o for o in off if o != 0)
        if first > 0:
            half.append(off)
    return half


def periodic_label(b, connectivity=None):
    """Label the solid pieces of a binary array whose faces wrap around (nlfiber/topology.py:8-46). Returns
    (labels, n): labels 1..n for the pieces, 0 on empty voxels."""
    b = np.asarray(b, dtype=bool)
    d = b.ndim
    idx = -np.ones(b.shape, dtype=np.int64)
    solid = np.flatnonzero(b)
    idx.flat[solid] = np.arange(solid.size)
    if solid.size == 0:
        return np.zeros(b.shape, dtype=np.int64), 0
    rows, cols = [], []
    for off in _offsets(d, connectivity):
        nb = np.roll(idx, shift=tuple(-o for o in off), axis=tuple(range(d)))  # index of the voxel at x + off
        both = (idx >= 0) & (nb >= 0)
        rows.append(idx[both])
        cols.append(nb[both])
    r = np.concatenate(rows) if rows else np.zeros(0, dtype=np.int64)
    c = np.concatenate(cols) if cols else np.zeros(0, dtype=np.int64)
    g = coo_matrix((np.ones(r.size, dtype=np.int8), (r, c)), shape=(solid.size, solid.size))
    n, comp = connected_components(g, directed=False)
    # number the components by their first voxel in C order
    first = np.full(n, solid.size, dtype=np.int64)
    np.minimum.at(first, comp, np.arange(solid.size))
    order = np.argsort(first)
    rank = np.empty(n, dtype=np.int64)
    rank[order] = np.arange(1, n + 1)
    labels = np.zeros(b.shape, dtype=np.int64)
    labels.flat[solid] = rank[comp]
    return labels, int(n)


def spanning(labels, k):
    """Whether component k has a voxel in each plane normal to each axis, i.e. it runs through the cell in all
    directions and can transmit force between opposite faces (nlfiber/topology.py:49-56)."""
    m = np.asarray(labels) == k
    for ax in range(m.ndim):
        other = tuple(a for a in range(m.ndim) if a != ax)
        if not m.any(axis=other).all():
            return False
    return True


def remove_floating(b):
    """The solid voxels of the spanning components only (nlfiber/topology.py:59-65)."""
    lab, n = periodic_label(b)
    keep = np.zeros(n + 1, dtype=bool)
    for k in range(1, n + 1):
        keep[k] = spanning(lab, k)
    return keep[lab]


def dilate_periodic(b):
    """Periodic dilation by the 26-neighbourhood (3x3x3 cube; 3x3 in 2D), nlfiber/density.py:124-129."""
    return ndimage.maximum_filter(np.asarray(b, dtype=bool), size=3, mode="wrap")


def upsample_periodic(c, f):
    """Periodic multilinear resampling of a cell-centred field by the factor f, axis by axis
    (nlfiber/density.py:132-148): sample j of the new axis of m = round(n f) points sits at
    x = (j + 1/2) n / m - 1/2 in old voxel units, between old samples floor(x) and floor(x) + 1 (mod n)."""
    out = np.asarray(c, dtype=np.float64)
    for ax in range(out.ndim):
        n = out.shape[ax]
        m = int(round(n * f))
        x = (np.arange(m) + 0.5) * n / m - 0.5
        i0 = np.floor(x).astype(np.int64)
        t = (x - i0).reshape([m if a == ax else 1 for a in range(out.ndim)])
        lo = np.take(out, np.mod(i0, n), axis=ax)
        hi = np.take(out, np.mod(i0 + 1, n), axis=ax)
        out = (1.0 - t) * lo + t * hi
    return out
