"""Synthetic unit cells for the benchmark workloads (input preparation, host numpy).

A cell is a periodic density rho on an (n, n, n) voxel grid with full cubic (Oh) symmetry: mirror
planes at the faces and mid-planes (voxel-centred reflection i -> n-1-i) and invariance under axis
permutations (nlfiber/symmetry.py:1-12 describes the same symmetry group on the octant).

Recipe (shared by every workload):
  1. pre-density phi at voxel centres x_i = 2 pi (i + 1/2) / n;
  2. periodic Gaussian filter of standard deviation r voxels (normalised kernel, applied by FFT);
  3. scale to unit standard deviation;
  4. exact symmetrisation: every Oh orbit takes the value of its canonical voxel, so thresholds never
     split an orbit (same purpose as nlfiber/symmetry.py:61-66);
  5. rho = sigmoid(beta (u - eta)) with eta solved for mean(rho) = v exactly by 200 bisection steps
     and 3 Newton steps (nlfiber/density.py:43-72).

The reference filters with a normalised periodic hat (cone) kernel max(0, r - d)
(nlfiber/density.py:23-31); `filter_kernel="hat"` selects it, the default follows the workload
text (Gaussian). Either way the solver only sees rho, which is stored with every fixture.
"""
import itertools

import numpy as np

OH_PERMS = list(itertools.permutations(range(3)))


def voxel_centres(n):
    return 2.0 * np.pi * (np.arange(n) + 0.5) / n


def schwarz_p(n):
    """phi = -(cos X + cos Y + cos Z) at voxel centres (tests/helpers.py:8-11 on the full cell)."""
    x = voxel_centres(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return -(np.cos(X) + np.cos(Y) + np.cos(Z))


def oh_average(f):
    """Average of a cubic array over the 48 maps (6 axis permutations x 8 voxel-centred flips)."""
    out = np.zeros_like(f, dtype=np.float64)
    for p in OH_PERMS:
        t = np.transpose(f, p)
        for flips in itertools.product((False, True), repeat=3):
            g = t
            for ax, fl in enumerate(flips):
                if fl:
                    g = np.flip(g, ax)
            out += g
    return out / 48.0


def random_fourier_phi(n, rng, kmax=2.0):
    """Sum of Fourier modes 0 < |k| <= kmax with N(0,1) cosine and sine coefficients, averaged over
    the 48 Oh maps. a cos(k.x) + b sin(k.x) = Re((a - i b) e^{i k.x}), and e^{i k.x} is the outer
    product of three 1-D exponentials, so the sum is three small contractions over a 5^3 coefficient
    grid."""
    x = voxel_centres(n)
    ks = [k for k in itertools.product(range(-2, 3), repeat=3) if 0 < np.dot(k, k) <= kmax * kmax]
    coef = np.zeros((5, 5, 5), np.complex128)
    for k in ks:
        a, b = rng.standard_normal(2)
        coef[k[0] + 2, k[1] + 2, k[2] + 2] += a - 1j * b
    e = np.exp(1j * np.outer(x, np.arange(-2, 3)))  # (n, 5)
    phi = np.einsum("abc,ia,jb,kc->ijk", coef, e, e, e, optimize=True).real
    return oh_average(phi)


def gaussian_kernel_fft(n, r):
    d = np.minimum(np.arange(n), n - np.arange(n)).astype(np.float64)
    D2 = d[:, None, None] ** 2 + d[None, :, None] ** 2 + d[None, None, :] ** 2
    k = np.exp(-D2 / (2.0 * r * r))
    k /= k.sum()
    return np.fft.rfftn(k)


def hat_kernel_fft(n, r):
    d = np.minimum(np.arange(n), n - np.arange(n)).astype(np.float64)
    dist = np.sqrt(d[:, None, None] ** 2 + d[None, :, None] ** 2 + d[None, None, :] ** 2)
    k = np.maximum(0.0, r - dist)
    if k.sum() == 0:
        k[0, 0, 0] = 1.0
    k /= k.sum()
    return np.fft.rfftn(k)


def periodic_filter(x, r, kind="gaussian"):
    n = x.shape[0]
    kf = gaussian_kernel_fft(n, r) if kind == "gaussian" else hat_kernel_fft(n, r)
    return np.fft.irfftn(np.fft.rfftn(x) * kf, s=x.shape, axes=tuple(range(x.ndim)))


_CANON = {}


def canonical_index(n):
    """Flat index of the canonical voxel of each voxel's Oh orbit: fold each axis to min(i, n-1-i),
    then sort the three folded indices."""
    if n not in _CANON:
        i = np.arange(n)
        f = np.minimum(i, n - 1 - i)
        F = np.stack(np.meshgrid(f, f, f, indexing="ij"), axis=-1)
        F = np.sort(F, axis=-1)
        _CANON[n] = (F[..., 0] * n + F[..., 1]) * n + F[..., 2]
    return _CANON[n]


def exact_symmetrize(u):
    u = oh_average(u)
    n = u.shape[0]
    return u.reshape(-1)[canonical_index(n)]


def _sigmoid(t):
    return 0.5 * (1.0 + np.tanh(0.5 * t))


def solve_eta(u, v, beta):
    """eta with mean(sigmoid(beta (u - eta))) = v: 200 bisection steps, then 3 Newton steps
    (nlfiber/density.py:43-63)."""
    lo = u.min() - 40.0 / beta
    hi = u.max() + 40.0 / beta
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        old = (lo, hi)
        if np.mean(_sigmoid(beta * (u - mid))) > v:
            lo = mid
        else:
            hi = mid
        if (lo, hi) == old or lo == hi:
            break  # the bracket has stopped moving: the remaining steps would not change it
    eta = 0.5 * (lo + hi)
    for _ in range(3):
        s = _sigmoid(beta * (u - eta))
        dv = -np.mean(beta * s * (1 - s))
        eta = eta - (np.mean(s) - v) / (dv if dv != 0 else -1.0)
    return eta


def density(phi, r=1.5, beta=16.0, v=0.3, filter_kernel="gaussian"):
    """phi -> rho through filter, unit-std scaling, exact symmetrisation and the exact-volume
    threshold."""
    u = periodic_filter(np.asarray(phi, np.float64), r, filter_kernel)
    u = u / (np.std(u) + 1e-12)
    u = exact_symmetrize(u)
    eta = solve_eta(u, v, beta)
    return _sigmoid(beta * (u - eta))


def schwarz_cell(n=32, r=1.5, beta=16.0, v=0.3):
    """The 32^3 Schwarz-P cell: phi = -(cos X + cos Y + cos Z), Gaussian r = 1.5, beta = 16,
    volume 0.3."""
    return density(schwarz_p(n), r=r, beta=beta, v=v)


def random_cell(n, seed, r=1.5, beta=None, v=None, betas=(2.0, 4.0, 8.0, 16.0), vrange=(0.2, 0.5)):
    """Seeded random cubic-symmetric cell: Fourier modes |k| <= 2, beta drawn from `betas` and the
    volume fraction from U[vrange] unless given. Returns (rho, meta)."""
    rng = np.random.default_rng(seed)
    phi = random_fourier_phi(n, rng)
    if beta is None:
        beta = float(betas[rng.integers(len(betas))])
    if v is None:
        v = float(rng.uniform(*vrange))
    rho = density(phi, r=r, beta=beta, v=v)
    return rho, dict(n=n, seed=int(seed), r=r, beta=beta, v=v)


def stress_cell(seed, n=128):
    """High-contrast cell: volume 0.15, beta 16 (used with emin = 1e-6)."""
    return random_cell(n, seed, beta=16.0, v=0.15)
