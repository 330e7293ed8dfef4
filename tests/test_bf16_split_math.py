"""The arithmetic of the BF16-split contraction, restated on the CPU.

Each FP32 operand x is split as hi = bf16(x), lo = bf16(x - hi) (round to
nearest even), and a product a b is taken as ah bh + ah bl + al bh.  These
checks pin the precision claims the device path relies on: x - hi - lo stays
within 2^-16 |x| (hi is within 2^-8 |x|, lo within 2^-8 of x - hi), so the
three-product sum is within 2^-14 of a b relative, hundreds of times closer
than one BF16 product.
"""
import numpy as np


def bf16_rne(x: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), returned
    as float32 (the bit trick the kernel uses: add 0x7FFF plus the lowest kept
    bit, then truncate the low 16 bits)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r << 16).astype(np.uint32).view(np.float32)


def split(x):
    hi = bf16_rne(x)
    lo = bf16_rne(np.float32(x) - hi)
    return hi, lo


def test_rne_matches_exact_rounding():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(20000) * 10.0 ** rng.uniform(-8, 8, 20000)).astype(np.float32)
    r = bf16_rne(x).astype(np.float64)
    # the nearest bfloat16 neighbours: 8 significant bits
    e = np.floor(np.log2(np.abs(x.astype(np.float64))))
    ulp = 2.0 ** (e - 7)
    assert np.all(np.abs(r - x) <= 0.5 * ulp * (1 + 1e-12))
    # exactly representable values are fixed points
    y = bf16_rne(x)
    assert np.array_equal(bf16_rne(y), y)


def test_ties_round_to_even():
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: ties go to the even 1.0
    assert bf16_rne(np.float32(1.0 + 2.0 ** -8)) == np.float32(1.0)
    # 1 + 3 * 2^-8 lies between 1 + 2^-7 (odd) and 1 + 2^-6 (even): rounds up
    assert bf16_rne(np.float32(1.0 + 3 * 2.0 ** -8)) == np.float32(1.0 + 2.0 ** -6)


def test_split_and_product_error():
    rng = np.random.default_rng(2)
    a = rng.standard_normal(50000).astype(np.float32)
    b = rng.standard_normal(50000).astype(np.float32)
    ah, al = split(a)
    bh, bl = split(b)
    ad, bd = a.astype(np.float64), b.astype(np.float64)
    resid = np.abs(ad - ah.astype(np.float64) - al.astype(np.float64))
    assert np.all(resid <= 2.0 ** -16 * np.abs(ad))
    approx = (ah.astype(np.float64) * bh + ah.astype(np.float64) * bl
              + al.astype(np.float64) * bh)
    rel = np.abs(approx - ad * bd) / np.abs(ad * bd)
    assert np.max(rel) <= 2.0 ** -14
    # and it is far better than one bfloat16 product
    one = np.abs(ah.astype(np.float64) * bh - ad * bd) / np.abs(ad * bd)
    assert np.median(rel) < 1e-2 * np.median(one)
