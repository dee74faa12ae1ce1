"""The arithmetic of the BF16 split-precision contraction
(SHARP_HIP_MIXED_BF16X3, planck-levels_amd/csrc/legendre_mixed.hip: split3,
mfma_bf3), checked on the CPU with numpy bit operations:

* truncating an FP32 value to its top 16 bits three times gives pieces
  h, m, l with x = h + m + l exactly, each piece a BF16 value;
* the six piece products kept (hh', hm', mh', hl', mm', lh') are exact in
  FP32 and miss x*y by less than 2^-20 |x y| (the dropped ml', lm', ll':
  |m| < 2^-7 |x| and |l| < 2^-14 |x|), about 2^-24 in the rms;
* a 4096-term dot product formed the way the kernels form it (six products,
  32 terms per matrix instruction, FP32 accumulation) is as close to the
  FP64 result as the FP32 matrix-core form, while the two-piece form with
  three products (the plain "bf16x3" of the literature) misses the 1e-6
  bound of the mixed forms by an order of magnitude.
"""
import numpy as np


def trunc16(x):
    return (np.asarray(x, np.float32).view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)


def split3(x):
    x = np.asarray(x, np.float32)
    h = trunc16(x)
    r1 = (x - h).astype(np.float32)
    m = trunc16(r1)
    r2 = (r1 - m).astype(np.float32)
    return h, m, trunc16(r2)


def is_bf16(v):
    return np.all((np.asarray(v, np.float32).view(np.uint32) & np.uint32(0xFFFF)) == 0)


def test_split_is_exact_and_bf16():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(200000) * np.exp(rng.uniform(-60, 60, 200000))).astype(np.float32)
    h, m, l = split3(x)
    assert is_bf16(h) and is_bf16(m) and is_bf16(l)
    # exact: summed in float64, no rounding can hide a difference
    assert np.array_equal(h.astype(np.float64) + m + l, x.astype(np.float64))
    # each piece is at most 2^-7 of the one before
    nz = m != 0
    assert np.all(np.abs(m[nz]) <= np.abs(h[nz]) * 2.0 ** -7)
    nz = l != 0
    assert np.all(np.abs(l[nz]) <= np.abs(m[nz]) * 2.0 ** -7)


def test_six_products_bound():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(100000).astype(np.float32)
    y = rng.standard_normal(100000).astype(np.float32)
    xs, ys = split3(x), split3(y)
    pairs = [(0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0)]
    prods = [xs[i].astype(np.float64) * ys[j] for i, j in pairs]
    for p, (i, j) in zip(prods, pairs):
        # a product of two 8-bit significands is exact in FP32
        assert np.array_equal(p.astype(np.float32).astype(np.float64), p)
    six = np.sum(prods, axis=0)
    exact = x.astype(np.float64) * y
    rel = np.abs(six - exact) / np.abs(exact)
    assert np.all(rel < 2.0 ** -20)
    assert float(np.sqrt(np.mean(rel ** 2))) < 2.0 ** -23


def _dot_fp32_mfma(a, b, K=4):
    acc = np.float32(0)
    a32, b32 = a.astype(np.float32), b.astype(np.float32)
    for k in range(0, len(a), K):
        acc = np.float32(acc + np.float32(np.dot(a32[k:k + K].astype(np.float64), b32[k:k + K])))
    return float(acc)


def _dot_split(a, b, pairs, K=32):
    xs, ys = split3(a), split3(b)
    acc = np.float32(0)
    for k in range(0, len(a), K):
        grp = np.float32(0)
        for i, j in pairs:
            grp = np.float32(grp + np.float32(np.dot(xs[i][k:k + K].astype(np.float64),
                                                     ys[j][k:k + K])))
        acc = np.float32(acc + grp)
    return float(acc)


def test_dot_products_like_the_transforms():
    """lambda-like oscillating rows against a_lm with a red spectrum."""
    rng = np.random.default_rng(2)
    L = 4096
    l = np.arange(L)
    six = [(2, 0), (1, 1), (0, 2), (1, 0), (0, 1), (0, 0)]
    three = [(1, 0), (0, 1), (0, 0)]
    e32, e6, e3 = [], [], []
    for _ in range(12):
        lam = np.cos(l * rng.uniform(0.1, 3.0) + rng.uniform(0, 6)) * \
            np.sqrt((2 * l + 1) / (4 * np.pi)) * rng.uniform(0.5, 1.5, L)
        alm = rng.standard_normal(L) / (l + 1.0)
        ref = float(np.dot(lam, alm))
        scale = float(np.sqrt(np.sum((lam * alm) ** 2)))
        e32.append(abs(_dot_fp32_mfma(lam, alm) - ref) / scale)
        e6.append(abs(_dot_split(lam, alm, six) - ref) / scale)
        # the two-piece form: pieces h, m only
        e3.append(abs(_dot_split(lam, alm, three) - ref) / scale)
    rms = lambda v: float(np.sqrt(np.mean(np.square(v))))  # noqa: E731
    assert rms(e6) < 1.5 * rms(e32), (rms(e6), rms(e32))
    assert rms(e6) < 1e-6
    assert rms(e3) > 1e-5
