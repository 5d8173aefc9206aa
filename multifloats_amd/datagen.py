"""Seeded synthetic inputs for the DD hot path (numpy, host side).

Arrays of DD values are numpy float64 arrays with a trailing axis of 2
(hi, lo), i.e. the float64x2 memory layout; matrices are column-major
(Fortran order) as in the reference's matmul convention
(reference README.md:227-245).

Generators (one per workload):
  config0        hi ~ U(-1,1); lo = hi * 2^-53 * U(-1,1), then Fast2Sum so
                 |lo| <= ulp(hi)/2.
  spread(e)      hi = s * 2^U(-e,e) * u, s = +-1, u ~ U[0.5,1); lo as config0.
  exp/log/sin/erf inputs of the elementwise workload (see each function).
  cancel_pairs   dot/sum inputs where 10% of adjacent term pairs nearly cancel.
"""
import numpy as np

_SPLIT = 134217729.0  # 2^27 + 1 (Veltkamp)


def _fast_two_sum(a, b):
    s = a + b
    return s, b - (s - a)


def _two_sum(a, b):
    s = a + b
    ap = s - b
    bp = s - ap
    return s, (a - ap) + (b - bp)


def _split(a):
    c = _SPLIT * a
    h = c - (c - a)
    return h, a - h


def two_prod(a, b):
    """Exact product a*b = p + e (Dekker), valid away from overflow/underflow."""
    p = a * b
    ah, al = _split(a)
    bh, bl = _split(b)
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl
    return p, e


def dd_mul_np(x, y):
    """DD product with the reference operator* order (numpy, finite inputs)."""
    p, e = two_prod(x[..., 0], y[..., 0])
    e = e + (x[..., 0] * y[..., 1] + x[..., 1] * y[..., 0])
    h, lo = _fast_two_sum(p, e)
    return np.stack([h, lo], axis=-1)


def _with_lo(rng, hi):
    lo = hi * 2.0 ** -53 * rng.uniform(-1.0, 1.0, size=hi.shape)
    h, lo = _fast_two_sum(hi, lo)
    return np.stack([h, lo], axis=-1)


def config0(rng, shape):
    shape = (shape,) if np.isscalar(shape) else tuple(shape)
    return _with_lo(rng, rng.uniform(-1.0, 1.0, size=shape))


def spread(rng, shape, e):
    shape = (shape,) if np.isscalar(shape) else tuple(shape)
    s = np.where(rng.random(shape) < 0.5, -1.0, 1.0)
    u = rng.uniform(0.5, 1.0, size=shape)
    hi = s * np.exp2(rng.uniform(-e, e, size=shape)) * u
    return _with_lo(rng, hi)


def as_colmajor(a):
    """View/copy an (m, n, 2) DD matrix in the column-major float64x2 layout:
    element (i, j) is the 16-byte pair at offset 16*(i + m*j), i.e. strides
    (16, 16*m, 8). (Fortran order of the 3-d array would split the limbs.)"""
    a = np.asarray(a, dtype=np.float64)
    return np.ascontiguousarray(a.transpose(1, 0, 2)).transpose(1, 0, 2)


def is_colmajor(a):
    if a.ndim != 3 or a.shape[2] != 2 or a.dtype != np.float64:
        return False
    if a.size == 0:
        return True
    if a.strides[2] != 8:
        return False
    m, n = a.shape[:2]
    return (m <= 1 or a.strides[0] == 16) and (n <= 1 or a.strides[1] == 16 * m)


def matrix(rng, m, n, e=None):
    """m x n DD matrix, (m, n, 2) array in the column-major float64x2 layout."""
    a = config0(rng, (n, m)) if e is None else spread(rng, (n, m), e)
    return a.transpose(1, 0, 2)


def flat(a):
    """(m*n, 2) array of a column-major DD matrix's elements in memory order."""
    return as_colmajor(a).transpose(1, 0, 2).reshape(-1, 2)


def unflat(f, m, n):
    """Inverse of flat()."""
    return np.asarray(f, dtype=np.float64).reshape(n, m, 2).transpose(1, 0, 2)


def exp_inputs(rng, n):
    return _with_lo(rng, rng.uniform(-700.0, 700.0, size=n))


def log_inputs(rng, n):
    return _with_lo(rng, 10.0 ** rng.uniform(-300.0, 300.0, size=n))


def erf_inputs(rng, n):
    return _with_lo(rng, rng.uniform(-6.0, 6.0, size=n))


def exp10_inputs(rng, n):
    """x in [-300, 300]: 10^x stays a normal double."""
    return _with_lo(rng, rng.uniform(-300.0, 300.0, size=n))


def acosh_inputs(rng, n):
    """x in [1, 1e3], log-spaced above 1 (acosh's domain)."""
    return _with_lo(rng, 1.0 + 10.0 ** rng.uniform(-4.0, 3.0, size=n))


def atanh_inputs(rng, n):
    """x in (-1, 1), within 1e-4 of neither end."""
    return _with_lo(rng, rng.uniform(-0.9999, 0.9999, size=n))


def sin_inputs(rng, n, near_frac=0.01):
    """hi ~ U[-1e4, 1e4]; a fraction of points sits within 1e-12 of k*pi/2.

    k*pi/2 is formed in double-double from a three-limb pi/2 (exact products
    of the integer k with each limb), so the planted points really are within
    1e-12 of a multiple of pi/2, where range reduction cancels the most.
    """
    hi = rng.uniform(-1e4, 1e4, size=n)
    x = _with_lo(rng, hi)
    nn = int(n * near_frac)
    if nn:
        idx = rng.choice(n, size=nn, replace=False)
        k = np.round(rng.uniform(-6366.0, 6366.0, size=nn))
        c1, c2 = 1.5707963267948966, 6.123233995736766e-17
        p1, e1 = two_prod(k, np.full(nn, c1))
        p2 = k * c2
        h, lo = _two_sum(p1, e1 + p2)
        h, lo = _fast_two_sum(h, lo)
        d = rng.uniform(-1e-12, 1e-12, size=nn)
        h2, l2 = _two_sum(h, d)
        h2, l2 = _fast_two_sum(h2, l2 + lo)
        x[idx, 0] = h2
        x[idx, 1] = l2
    return x


def cancel_pairs(rng, n, frac=0.10):
    """(a, b) with a, b ~ config0; in 10% of adjacent pairs the terms nearly
    cancel: a[2i+1] = a[2i], b[2i+1] = -b[2i] * (1 + d), d ~ U(-1e-20, 1e-20)."""
    a = config0(rng, n)
    b = config0(rng, n)
    npairs = n // 2
    sel = np.nonzero(rng.random(npairs) < frac)[0]
    i0, i1 = 2 * sel, 2 * sel + 1
    a[i1] = a[i0]
    d = np.stack([np.ones(sel.size), rng.uniform(-1e-20, 1e-20, size=sel.size)], axis=-1)
    b[i1] = -dd_mul_np(b[i0], d)
    return a, b


def sum_inputs(rng, n, frac=0.10):
    """x ~ config0 with 10% of adjacent pairs x[2i+1] = -x[2i] * (1 + d)."""
    x = config0(rng, n)
    npairs = n // 2
    sel = np.nonzero(rng.random(npairs) < frac)[0]
    d = np.stack([np.ones(sel.size), rng.uniform(-1e-20, 1e-20, size=sel.size)], axis=-1)
    x[2 * sel + 1] = -dd_mul_np(x[2 * sel], d)
    return x


def batched_sizes(rng, ngemm=10000, ndot=50000, lo=8, hi=512, dlo=16, dhi=65536):
    """Problem sizes of the batching workload: m, n, k of each GEMM and the
    length of each dot are log-uniform integers in [lo, hi] / [dlo, dhi]."""
    def logu(a, b, size):
        return np.clip(np.round(np.exp(rng.uniform(np.log(a), np.log(b), size))), a, b).astype(np.int64)
    return logu(lo, hi, (ngemm, 3)), logu(dlo, dhi, ndot)


def fill_config0_torch(torch, t, seed):
    """Fill a float64 CUDA tensor viewed as (..., 2) with config-0 DD values
    on the device (hi ~ U(-1,1), lo = hi 2^-53 U(-1,1), then Fast2Sum)."""
    g = torch.Generator(device=t.device)
    g.manual_seed(seed)
    v = t.view(-1, 2)
    hi = torch.rand(v.shape[0], generator=g, device=t.device, dtype=torch.float64) * 2.0 - 1.0
    lo = hi * (2.0 ** -53)
    lo = lo * (torch.rand(v.shape[0], generator=g, device=t.device, dtype=torch.float64) * 2.0 - 1.0)
    s = hi + lo
    v[:, 1] = lo - (s - hi)
    v[:, 0] = s
