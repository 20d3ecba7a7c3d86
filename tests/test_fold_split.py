"""The folded first pass carries a float32 norm as three bfloat16 parts whose
sum is the norm exactly (kernels_prep.hip split3_bf16).  Checked here on the
CPU with a bit-level bfloat16 round-to-nearest-even emulation over random
floats of every magnitude and the awkward cases (ties, carries into the next
binade, values just below a power of two).  The argument needs the remainders
to stay normal floats and the first part finite: 2^-100 <= |v| < 2^127 (below,
the parts' sum is within the subnormal spacing 2^-149 of v)."""
import numpy as np


def _bf16_rne(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def _split3(v):
    v = np.asarray(v, np.float32)
    h1 = _bf16_rne(v)
    r1 = (v - h1).astype(np.float32)
    h2 = _bf16_rne(r1)
    r2 = (r1 - h2).astype(np.float32)
    h3 = _bf16_rne(r2)
    return h1, h2, h3, r2


def test_rne_emulation_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1
    assert _bf16_rne(np.float32(1 + 2**-8)) == 1.0
    # 1 + 3 * 2^-8 is a tie between 1 + 2^-7 and 1 + 2^-6: even -> 1 + 2^-6
    assert _bf16_rne(np.float32(1 + 3 * 2**-8)) == np.float32(1 + 2**-6)
    assert _bf16_rne(np.float32(1 + 2**-8 + 2**-20)) == np.float32(1 + 2**-7)
    assert _bf16_rne(np.float32(-2.5)) == -2.5


def test_three_parts_sum_exactly():
    rng = np.random.default_rng(0)
    mant = rng.integers(0, 1 << 23, 200_000, dtype=np.uint64)
    expo = rng.integers(27, 254, 200_000, dtype=np.uint64)  # 2^-100 <= |v| < 2^127
    bits = ((expo << 23) | mant).astype(np.uint32)
    v = bits.view(np.float32)
    special = np.array([1.0, 1 + 2**-8, 1 + 3 * 2**-8, 2 - 2**-23, 2 - 2**-9, 1.5 + 2**-9,
                        1.7e38, 2.0**-100, 12345.678, 0.0], np.float32)
    # magnitudes that norms take (squared norms of vectors of any scale)
    norms = (rng.random(100_000) * 10.0 ** rng.uniform(-6, 8, 100_000)).astype(np.float32)
    v = np.concatenate([v, special, norms])
    v = v[np.isfinite(v)]
    h1, h2, h3, r2 = _split3(v)
    # each remainder is exact in float32 and the last one is a bfloat16
    assert (h3 == r2).all()
    total = h1.astype(np.float64) + h2.astype(np.float64) + h3.astype(np.float64)
    assert (total == v.astype(np.float64)).all()
