"""Seeded synthetic systems for the solver workloads (host side, numpy).

All matrices are returned as 0-based CSR triples ``(rowptr int32, colidx int32,
val float64)`` with sorted columns and the diagonal present, unknowns numbered
lexicographically (x fastest).  The reference ships no generators; the
definitions below are this package's own and are deterministic for a given
numpy version (PCG64 streams from ``numpy.random.default_rng``).

* ``poisson2d(N)``        -- 5-point Laplacian, Dirichlet boundary eliminated,
                             stencil [4, -1, -1, -1, -1].
* ``convdiff3d(N, pe)``   -- 7-point convection-diffusion, D = 1, velocity
                             direction (1, 0.5, 0.25) scaled so the largest
                             cell Peclet number |v_k| h / (2D) is ``pe``;
                             central differences, scaled by h^2:
                             diag 6, neighbour along +k: -(1 - P_k),
                             along -k: -(1 + P_k).
* ``lognormal27(N, seed)``-- 27-point stencil, off-diagonal -exp(N(0,1)) drawn
                             once per undirected edge (symmetric pattern and
                             values), diagonal = dom * sum |off-diagonals|.
* ``vardiff2d(kappa)``    -- 5-point variable-coefficient diffusion: interior
                             face coefficient = harmonic mean of the two cell
                             kappas, boundary faces use the cell's kappa;
                             kappa = 1 gives ``poisson2d``.
* ``grf_kappa(N, ell, seed)`` -- kappa = exp(g), g a Gaussian random field with
                             squared-exponential covariance of length ``ell`` on
                             the unit square, by spectral synthesis (white noise
                             filtered in Fourier space by the square root of the
                             kernel's spectral density, then scaled to unit
                             sample variance).
"""
import numpy as np


def _assemble(n, cols, vals, mask):
    """Build CSR from per-row candidate entries (already sorted by column)."""
    cnt = mask.sum(axis=1)
    rowptr = np.zeros(n + 1, np.int64)
    np.cumsum(cnt, out=rowptr[1:])
    colidx = cols[mask].astype(np.int32)
    val = vals[mask].astype(np.float64)
    return rowptr.astype(np.int32), colidx, val


def _grid_offsets(dims, offs):
    """For a lexicographic grid, return (cols, mask) for stencil offsets."""
    shape = tuple(reversed(dims))  # (nz, ny, nx) or (ny, nx)
    n = int(np.prod(dims))
    idx = np.arange(n, dtype=np.int64)
    coords = np.unravel_index(idx, shape)  # slowest first
    coords = list(reversed(coords))       # x, y, z
    cols = np.empty((n, len(offs)), np.int64)
    mask = np.ones((n, len(offs)), bool)
    strides = [1]
    for d in dims[:-1]:
        strides.append(strides[-1] * d)
    for k, off in enumerate(offs):
        lin = 0
        for d, o in enumerate(off):
            c = coords[d] + o
            mask[:, k] &= (c >= 0) & (c < dims[d])
            lin += o * strides[d]
        cols[:, k] = idx + lin
    cols[~mask] = 0
    return cols, mask


def poisson2d(N):
    offs = [(0, -1), (-1, 0), (0, 0), (1, 0), (0, 1)]
    cols, mask = _grid_offsets((N, N), offs)
    vals = np.tile(np.array([-1.0, -1.0, 4.0, -1.0, -1.0]), (N * N, 1))
    return _assemble(N * N, cols, vals, mask)


def convdiff3d(N, pe=0.5, direction=(1.0, 0.5, 0.25)):
    d = np.asarray(direction, np.float64)
    P = pe * d / np.abs(d).max()  # cell Peclet per axis (x, y, z)
    offs = [(0, 0, -1), (0, -1, 0), (-1, 0, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    coef = np.array([-(1 + P[2]), -(1 + P[1]), -(1 + P[0]), 6.0,
                     -(1 - P[0]), -(1 - P[1]), -(1 - P[2])])
    cols, mask = _grid_offsets((N, N, N), offs)
    vals = np.broadcast_to(coef, cols.shape)
    return _assemble(N ** 3, cols, vals, mask)


def lognormal27(N, seed=2, dom=1.05):
    n = N ** 3
    offs = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    cols, mask = _grid_offsets((N, N, N), offs)
    fwd = [k for k, o in enumerate(offs) if (o[2], o[1], o[0]) > (0, 0, 0)]
    rng = np.random.default_rng(seed)
    W = np.exp(rng.standard_normal((len(fwd), n)))
    vals = np.zeros((n, 27))
    strides = np.array([1, N, N * N])
    for m, k in enumerate(fwd):
        o = np.array(offs[k])
        lin = int((o * strides).sum())
        kb = offs.index(tuple(-o))
        # forward edge (i, i+lin) carries W[m, i]; the backward entry of row j
        # = i + lin refers to the same edge.
        vals[:, k] = -W[m]
        src = np.arange(n) - lin
        ok = mask[:, kb]
        vals[ok, kb] = -W[m, src[ok]]
    vals[~mask] = 0.0
    diag = dom * np.abs(vals).sum(axis=1)
    vals[:, 13] = diag
    return _assemble(n, cols, vals, mask)


def vardiff2d(kappa):
    ny, nx = kappa.shape
    n = nx * ny
    k = kappa.reshape(-1)
    offs = [(0, -1), (-1, 0), (0, 0), (1, 0), (0, 1)]
    cols, mask = _grid_offsets((nx, ny), offs)
    vals = np.zeros((n, 5))
    diag = np.zeros(n)
    for j in (0, 1, 3, 4):
        kn = k[cols[:, j]]
        face = np.where(mask[:, j], 2.0 * k * kn / (k + kn), k)
        vals[:, j] = np.where(mask[:, j], -face, 0.0)
        diag += face
    vals[:, 2] = diag
    return _assemble(n, cols, vals, mask)


def grf_kappa(N, ell=0.1, seed=4):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((N, N))
    kx = np.fft.fftfreq(N, d=1.0 / N) * 2 * np.pi
    ky = np.fft.rfftfreq(N, d=1.0 / N) * 2 * np.pi
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    spec = np.exp(-0.5 * ell * ell * k2)
    g = np.fft.irfft2(np.fft.rfft2(w) * np.sqrt(spec), s=(N, N))
    g = (g - g.mean()) / g.std()
    return np.exp(g)


def spmv(A, x):
    rp, ci, v = A
    n = len(rp) - 1
    lens = np.diff(rp)
    rows = np.repeat(np.arange(n), lens)
    return np.bincount(rows, weights=v * x[ci], minlength=n)


def manufactured(A, seed):
    """x* ~ U[-1, 1] from ``default_rng(seed)`` and b = A x*."""
    n = len(A[0]) - 1
    xs = np.random.default_rng(seed).uniform(-1.0, 1.0, n)
    return xs, spmv(A, xs)


def batch_systems(count=512, smin=32, smax=160, seed=3, sigma=0.5):
    """Independent 2D variable-coefficient diffusion systems of uneven size."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(smin, smax + 1, size=count)
    systems = []
    for s in sizes:
        kappa = np.exp(sigma * rng.standard_normal((s, s)))
        systems.append(vardiff2d(kappa))
    return sizes, systems


# Named workloads: (builder, rhs seed)
def workload(name):
    if name == "poisson2d_256":
        A = poisson2d(256)
        return A, manufactured(A, 0)
    if name == "convdiff3d_128":
        A = convdiff3d(128)
        return A, manufactured(A, 1)
    if name == "lognormal27_100":
        A = lognormal27(100, seed=2)
        return A, manufactured(A, [2, 1])
    if name == "grf2d_512":
        A = vardiff2d(grf_kappa(512, 0.1, 4))
        return A, manufactured(A, [4, 1])
    raise KeyError(name)
