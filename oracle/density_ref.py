"""Float64 numpy restatement of the reference's density map and cubic symmetry helpers (test infrastructure).

TEST INFRASTRUCTURE ONLY: the GPU tests compare the library's density-map kernels and their
vector-Jacobian products against these functions; the product never imports them.

Follows nlfiber/symmetry.py:26-102 (sym_perm, mirror, fold, exact_symmetrize, free entries) and
nlfiber/density.py:23-91 (hat filter, periodic convolution, eta by bisection + Newton, implicit-derivative
threshold, DensityMap.smooth / rho), in three dimensions. `rho_vjp` is the hand-derived adjoint of
`rho` (pinned by central differences in tests/test_oracle.py). `binary` restates DensityMap.binary
(nlfiber/density.py:96-148) with the periodic components of nlfiber/topology.py:8-65 found by a plain
breadth-first search over the 26-neighbourhood (`label_bfs`, pinned on hand-made cells in tests/test_oracle.py).
"""
from collections import deque
import itertools

import numpy as np

PERMS = list(itertools.permutations(range(3)))


def sym_perm(x):
    """Average over the 6 axis permutations (nlfiber/symmetry.py:26-35), summed in itertools order."""
    out = 0.0
    for p in PERMS:
        out = out + np.transpose(x, p)
    return out / len(PERMS)


def mirror(oct_):
    """Octant (n_o)^3 -> full cell (2 n_o)^3 (nlfiber/symmetry.py:38-43)."""
    x = oct_
    for ax in range(3):
        x = np.concatenate([x, np.flip(x, ax)], axis=ax)
    return x


def fold(full):
    """Full cell -> octant, averaging the mirror images (nlfiber/symmetry.py:46-54)."""
    x = full
    for ax in range(3):
        n = x.shape[ax] // 2
        a = np.take(x, np.arange(n), axis=ax)
        b = np.flip(np.take(x, np.arange(n, 2 * n), axis=ax), ax)
        x = 0.5 * (a + b)
    return x


def free_index(n_o):
    return np.array(list(itertools.combinations_with_replacement(range(n_o), 3)))


def exact_symmetrize(full):
    """Orbit values copied from the sorted representative (nlfiber/symmetry.py:62-68, 82-102)."""
    oct_ = sym_perm(fold(full))
    n_o = oct_.shape[0]
    g = np.stack(np.meshgrid(*[np.arange(n_o)] * 3, indexing="ij"), axis=-1)
    s = np.sort(g, axis=-1)
    return mirror(oct_[s[..., 0], s[..., 1], s[..., 2]])


def group_average(full):
    """Mean over the 48 maps (8 mirrors x 6 permutations) of a full-cell field: the orthogonal projection
    onto cubic-symmetric fields; exact_symmetrize computes it, and it is its own adjoint."""
    out = 0.0
    for p in PERMS:
        t = np.transpose(full, p)
        for f in itertools.product((False, True), repeat=3):
            out = out + np.flip(t, [a for a in range(3) if f[a]]) if any(f) else out + t
    return out / 48.0


def hat_kernel(shape, r):
    """Normalised periodic hat (cone) kernel max(0, r - dist) (nlfiber/density.py:23-31), spatial form."""
    grids = np.meshgrid(*[np.minimum(np.arange(n), n - np.arange(n)) for n in shape], indexing="ij")
    dist = np.sqrt(sum(g.astype(np.float64) ** 2 for g in grids))
    k = np.maximum(0.0, r - dist)
    if k.sum() == 0:
        k[(0,) * len(shape)] = 1.0
    return k / k.sum()


def filt(x, r):
    """Periodic convolution with the hat kernel through rfftn (nlfiber/density.py:34-36)."""
    kf = np.fft.rfftn(hat_kernel(x.shape, r))
    return np.fft.irfftn(np.fft.rfftn(x) * kf, s=x.shape, axes=tuple(range(x.ndim)))


def sigmoid(t):
    return 0.5 * (1.0 + np.tanh(0.5 * t))


def solve_eta(u, v, beta):
    """mean(sigmoid(beta (u - eta))) = v: 200 bisection steps, then 3 Newton steps (nlfiber/density.py:43-63)."""
    lo = u.min() - 40.0 / beta
    hi = u.max() + 40.0 / beta
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if np.mean(sigmoid(beta * (u - mid))) > v:
            lo = mid
        else:
            hi = mid
    eta = 0.5 * (lo + hi)
    for _ in range(3):
        s = sigmoid(beta * (u - eta))
        dv = -np.mean(beta * s * (1 - s))
        eta = eta - (np.mean(s) - v) / (dv if dv != 0 else -1.0)
    return eta


def smooth(phi, r):
    """u = exact_symmetrize(F mirror(sym_perm(phi)) / std) (nlfiber/density.py:80-84)."""
    f = filt(mirror(sym_perm(phi)), r)
    u = f / (np.std(f) + 1e-12)
    return exact_symmetrize(u)


def rho(phi, beta, r, v):
    """H_beta(u - eta) with mean rho = v (nlfiber/density.py:66-72, 86-87)."""
    u = smooth(phi, r)
    return sigmoid(beta * (u - solve_eta(u, v, beta)))


def rho_vjp(phi, beta, r, v, rbar):
    """phi-gradient of <rbar, rho(phi)> for rows rbar (k, n, n, n): the adjoint of every step, with the implicit
    eta derivative of nlfiber/density.py:66-72 (d eta = d.du / sum d, d = s (1 - s))."""
    f = filt(mirror(sym_perm(phi)), r)
    m, sd = f.mean(), np.std(f)
    us = f / (sd + 1e-12)
    u = exact_symmetrize(us)
    eta = solve_eta(u, v, beta)
    s = sigmoid(beta * (u - eta))
    d = s * (1 - s)
    out = []
    for rb in np.asarray(rbar).reshape((-1,) + u.shape):
        ubar = beta * d * (rb - np.sum(rb * d) / np.sum(d))
        usbar = group_average(ubar)
        fbar = usbar / (sd + 1e-12) - np.sum(usbar * f) / (sd + 1e-12) ** 2 * (f - m) / (f.size * sd)
        mbar = filt(fbar, r)  # the hat kernel is even: the adjoint convolution is the same
        out.append(sym_perm(sum_mirror(mbar)))
    return np.stack(out)


def sum_mirror(full):
    """Adjoint of `mirror`: the octant receives the sum of its 8 images."""
    n_o = full.shape[0] // 2
    x = full
    for ax in range(3):
        a = np.take(x, np.arange(n_o), axis=ax)
        b = np.flip(np.take(x, np.arange(n_o, 2 * n_o), axis=ax), ax)
        x = a + b
    return x


# --- binary cells: nlfiber/density.py:96-148, nlfiber/topology.py:8-65 ---------------------------------------
def label_bfs(b, face_only=False):
    """Periodic connected components by breadth-first search (26-neighbourhood, or 6 with face_only), numbered
    1..n in the order of their first voxel in C order. Returns (labels, n)."""
    b = np.asarray(b, dtype=bool)
    shp = b.shape
    offs = [o for o in itertools.product((-1, 0, 1), repeat=b.ndim) if any(o)]
    if face_only:
        offs = [o for o in offs if sum(abs(t) for t in o) == 1]
    lab = np.zeros(shp, dtype=np.int64)
    n = 0
    for start in zip(*np.nonzero(b)):
        if lab[start]:
            continue
        n += 1
        lab[start] = n
        q = deque([start])
        while q:
            cur = q.popleft()
            for o in offs:
                nb = tuple((c + t) % m for c, t, m in zip(cur, o, shp))
                if b[nb] and not lab[nb]:
                    lab[nb] = n
                    q.append(nb)
    return lab, n


def spans(lab, k):
    m = lab == k
    for ax in range(m.ndim):
        if not np.all(m.any(axis=tuple(a for a in range(m.ndim) if a != ax))):
            return False
    return True


def drop_floating(b):
    lab, n = label_bfs(b)
    keep = np.array([False] + [spans(lab, k) for k in range(1, n + 1)])
    return keep[lab]


def upsample(c, f):
    """Periodic multilinear resampling, one axis at a time (nlfiber/density.py:132-148)."""
    out = np.asarray(c, dtype=np.float64)
    for ax in range(out.ndim):
        n = out.shape[ax]
        m = int(round(n * f))
        x = (np.arange(m) + 0.5) * n / m - 0.5
        i0 = np.floor(x).astype(int)
        w = (x - i0).reshape([m if a == ax else 1 for a in range(out.ndim)])
        out = (1 - w) * np.take(out, i0 % n, axis=ax) + w * np.take(out, (i0 + 1) % n, axis=ax)
    return out


def binary(phi, r, v, up=1):
    """Binary cell: smoothed u (upsampled and re-symmetrised when up > 1) above its (1 - v) quantile, floating
    pieces removed, the kept solid grown back to v through its void neighbours (highest u first)."""
    u = smooth(phi, r)
    if up > 1:
        u = exact_symmetrize(upsample(u, up))
    b = u > np.quantile(u, 1.0 - v)
    for _ in range(500):
        b = drop_floating(b)
        missing = v - b.mean()
        if missing * b.size < 1 or not b.any():
            break
        front = b.copy()
        for ax in range(b.ndim):
            front = front | np.roll(front, 1, axis=ax) | np.roll(front, -1, axis=ax)
        front &= ~b
        cand = np.sort(u[front])[::-1]
        if cand.size == 0:
            break
        k = min(max(int(round(missing * b.size)), 1), cand.size)
        b = b | (front & (u >= cand[k - 1]))
    return drop_floating(b)
