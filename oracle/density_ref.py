"""Float64 numpy restatement of the reference's density map and cubic symmetry helpers (test infrastructure).

TEST INFRASTRUCTURE ONLY: the GPU tests compare the library's density-map kernels and their
vector-Jacobian products against these functions; the product never imports them.

Follows nlfiber/symmetry.py:26-102 (sym_perm, mirror, fold, exact_symmetrize, free entries) and
nlfiber/density.py:23-91 (hat filter, periodic convolution, eta by bisection + Newton, implicit-derivative
threshold, DensityMap.smooth / rho), in three dimensions. `rho_vjp` is the hand-derived adjoint of
`rho` (pinned by central differences in tests/test_oracle.py).
"""
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
