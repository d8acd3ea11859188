"""Periodic connectivity and binarization helpers of the binary check (nlfiber/topology.py:8-65,
nlfiber/density.py:96-148) against hand-made cells and the oracle's breadth-first search."""
import itertools

import numpy as np
import pytest

from oracle import density_ref as DR


@pytest.fixture(scope="module")
def topo(pkg):
    return pkg.topology


def test_bfs_labels_on_hand_made_cells():
    b = np.zeros((6, 6, 6), bool)
    b[0, 0, :] = True  # a rod along z, closed through the periodic boundary
    lab, n = DR.label_bfs(b)
    assert n == 1 and not DR.spans(lab, 1)  # it does not meet every x slice
    b = np.zeros((6, 6, 6), bool)
    b[0, 0, 0] = b[5, 5, 5] = True  # corner neighbours through the periodic boundary
    assert DR.label_bfs(b)[1] == 1 and DR.label_bfs(b, face_only=True)[1] == 2
    b = np.zeros((5, 5, 5), bool)
    b[1, 1, 1] = b[3, 3, 3] = True
    assert DR.label_bfs(b)[1] == 2
    b = np.zeros((4, 4, 4), bool)
    b[:, 0, 0] = b[0, :, 0] = b[0, 0, :] = True  # three rods meeting at the origin: one spanning frame
    lab, n = DR.label_bfs(b)
    assert n == 1 and DR.spans(lab, 1)


@pytest.mark.parametrize("seed", range(6))
def test_periodic_label_matches_bfs(topo, seed):
    rng = np.random.default_rng(seed)
    shape = [(7, 6, 5), (8, 8, 8), (5, 9, 4), (6, 6), (3, 4, 5), (9, 9, 9)][seed]
    b = rng.random(shape) < (0.18 if len(shape) == 3 else 0.35)
    face = None if seed % 2 == 0 else np.zeros((3,) * len(shape), bool)
    if face is not None:
        for d in range(len(shape)):
            for o in (0, 2):
                idx = [1] * len(shape)
                idx[d] = o
                face[tuple(idx)] = True
    lab, n = topo.periodic_label(b, connectivity=face)
    if len(shape) == 3:
        ref, nr = DR.label_bfs(b, face_only=face is not None)
        assert n == nr and np.array_equal(lab, ref)  # same partition and the same numbering
    else:  # 2D: brute-force union of 8-neighbours (4 with the face stencil)
        offs = [o for o in itertools.product((-1, 0, 1), repeat=2) if any(o)]
        if face is not None:
            offs = [o for o in offs if abs(o[0]) + abs(o[1]) == 1]
        for i, j in zip(*np.nonzero(b)):
            for o in offs:
                q = ((i + o[0]) % shape[0], (j + o[1]) % shape[1])
                if b[q]:
                    assert lab[i, j] == lab[q]
        assert n == len(np.unique(lab[b]))
    assert lab[~b].max(initial=0) == 0


def test_remove_floating_keeps_spanning_pieces(topo):
    b = np.zeros((8, 8, 8), bool)
    b[:, 2, 2] = b[2, :, 2] = b[2, 2, :] = True  # spanning frame
    b[5:7, 5:7, 5:7] = True  # floating block
    kept = topo.remove_floating(b)
    assert kept.sum() == b.sum() - 8 and not kept[5:7, 5:7, 5:7].any()
    assert np.array_equal(kept, DR.drop_floating(b))


def test_upsample_and_dilation(topo):
    rng = np.random.default_rng(3)
    u = rng.standard_normal((6, 6, 6))
    assert np.array_equal(topo.upsample_periodic(u, 1), u)
    assert np.array_equal(topo.upsample_periodic(u, 2), DR.upsample(u, 2))
    assert np.array_equal(topo.upsample_periodic(u, 1.5), DR.upsample(u, 1.5))
    # multilinear: a field linear in x along a period-free direction is reproduced between samples
    x = np.arange(6, dtype=float)
    lin = np.broadcast_to(np.cos(2 * np.pi * (x + 0.5) / 6)[:, None, None], (6, 6, 6))
    up = topo.upsample_periodic(lin, 2)
    m = np.arange(12)
    xs = (m + 0.5) * 6 / 12 - 0.5
    i0 = np.floor(xs).astype(int)
    w = xs - i0
    want = (1 - w) * lin[i0 % 6, 0, 0] + w * lin[(i0 + 1) % 6, 0, 0]
    assert np.allclose(up[:, 3, 2], want, rtol=0, atol=1e-15)
    b = np.zeros((5, 5, 5), bool)
    b[0, 0, 0] = True
    d = topo.dilate_periodic(b)
    assert d.sum() == 27 and d[4, 4, 4] and d[1, 1, 1] and not d[2, 0, 0]


def test_host_exact_symmetrize_matches_oracle(pkg):
    rng = np.random.default_rng(5)
    u = rng.standard_normal((10, 10, 10))
    s = pkg.density.exact_symmetrize_host(u)
    assert np.array_equal(s, DR.exact_symmetrize(u))
    # every map of the cube group leaves it unchanged bitwise
    for p in itertools.permutations(range(3)):
        t = np.transpose(s, p)
        for f in itertools.product((False, True), repeat=3):
            ax = [a for a in range(3) if f[a]]
            assert np.array_equal(np.flip(t, ax) if ax else t, s)
