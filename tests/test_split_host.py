"""The operand split of the split-BF16 Coulomb forms (split_a in publicrelease_2020_amd/csrc/flosic_kernels.hip:
each bf16 piece the round-to-nearest-even of the remainder), restated bit for bit in numpy and checked against the
bounds the arithmetic fixes: each remainder is at most 2^-8 of the previous piece's value, three pieces carry an
FP32 value to 2^-24, and the six cross products the kernel keeps reproduce x*y to ~2^-22 while the three leading
ones alone (the 2-piece form) are off by ~2^-16.  A split that dropped a piece or truncated instead of rounding
fails one of these bounds."""
import numpy as np


def bf16_round(x):
    """Round-to-nearest-even of float32 values to bf16, returned as float32 ((__bf16)x for finite x)."""
    u = x.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def bf16_trunc(x):
    return (x.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)


def split(x, n, piece=bf16_round):
    """x (float32) -> n pieces and the final remainder, each remainder formed in FP32 as the kernel does."""
    out, r = [], x
    for _ in range(n):
        b = piece(r)
        out.append(b)
        r = (r - b).astype(np.float32)
    return out, r


def _samples(rng, n=200_000):
    # magnitudes of R over the Coulomb workloads span roughly 1e-30 .. 1e12; both signs
    e = rng.uniform(-100.0, 40.0, n)
    return (rng.choice([-1.0, 1.0], n) * rng.uniform(1.0, 2.0, n) * 2.0 ** e).astype(np.float32)


def test_rounded_pieces():
    x = _samples(np.random.default_rng(0))
    (b1, b2, b3), r3 = split(x, 3)
    for b in (b1, b2, b3):  # every piece is a bf16
        np.testing.assert_array_equal(b.view(np.uint32) & np.uint32(0xFFFF), 0)
    ax = np.abs(x.astype(np.float64))
    # each FP32 remainder is exact (Sterbenz) and at most half a bf16 ulp of the value it came from
    assert np.all(np.abs((x - b1).astype(np.float64)) <= ax * 2.0 ** -8)
    assert np.all(np.abs(b2.astype(np.float64)) <= ax * 2.0 ** -8 * (1 + 2.0 ** -8))
    # three pieces carry the value to 2^-24
    rest = np.abs(x.astype(np.float64) - b1 - b2.astype(np.float64) - b3.astype(np.float64))
    assert np.all(rest <= ax * 2.0 ** -24)
    # two pieces only to 2^-16
    (c1, c2), _ = split(x, 2)
    rest2 = np.abs(x.astype(np.float64) - c1 - c2.astype(np.float64))
    assert np.all(rest2 <= ax * 2.0 ** -16)
    assert rest2.max() / ax[np.argmax(rest2)] > 2.0 ** -20


def _product_error(xs, ys, x, y, nterms):
    xs = [v.astype(np.float64) for v in xs]
    ys = [v.astype(np.float64) for v in ys]
    terms = [(0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0)][:nterms]
    kept = sum(xs[i] * ys[j] for i, j in terms)
    exact = x.astype(np.float64) * y.astype(np.float64)
    return np.abs(kept - exact) / np.abs(exact)


def test_six_products_carry_the_product():
    rng = np.random.default_rng(2)
    x, y = _samples(rng, 50_000), _samples(rng, 50_000)
    xs, _ = split(x, 3)
    ys, _ = split(y, 3)
    assert _product_error(xs, ys, x, y, 6).max() < 2.0 ** -22
    # the three leading products alone (the 2-piece form) are ~2^-16 off: the bound tells the forms apart
    assert _product_error(xs, ys, x, y, 3).max() > 2.0 ** -18
    # truncated pieces shrink only by 2^-7: the same six products are then visibly worse
    xt, _ = split(x, 3, bf16_trunc)
    yt, _ = split(y, 3, bf16_trunc)
    assert _product_error(xt, yt, x, y, 6).max() > 2.0 ** -22
