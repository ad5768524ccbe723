"""The tile kernel's fixed-point CIC weights (particles.hip, fix_weight):
round(v * 2^44) for v in [0, 1] is taken from the low 52 bits of
v * 2^44 + 2^52 rounded once to double. Checked here with numpy's IEEE
arithmetic against round-half-even of the exact product, on random weights
of both particle precisions, on exact ties, and at 0 and 1."""
import numpy as np

K = 2.0 ** 44
MAGIC = 2.0 ** 52
MASK = np.uint64((1 << 52) - 1)


def fix_weight(v):
    m = np.asarray(v, np.float64) * K + MAGIC  # the product is exact, so this is one rounding, as the FMA
    return m.view(np.uint64) & MASK


def test_random_weights_both_precisions():
    rng = np.random.default_rng(1)
    for v in (rng.random(1_000_000), rng.random(1_000_000).astype(np.float32).astype(np.float64),
              rng.random(100_000) ** 8):  # small weights, where the product has a fractional part
        np.testing.assert_array_equal(fix_weight(v), np.rint(v * K).astype(np.uint64))


def test_ties_round_to_even_and_the_ends():
    # v * 2^44 = n + 1/2 exactly: n even rounds down, n odd rounds up
    n = np.arange(0, 1000, dtype=np.float64)
    v = (n + 0.5) / K
    got = fix_weight(v)
    np.testing.assert_array_equal(got, np.where(n % 2 == 0, n, n + 1).astype(np.uint64))
    np.testing.assert_array_equal(fix_weight(np.array([0.0, 1.0])), np.array([0, 1 << 44], np.uint64))


def test_eight_weights_sum_to_one_particle():
    """The eight weights of a particle sum to 2^44 units up to one unit per
    weight of rounding, so a window's charge is exact to 8 * 2^-44 of a
    particle's charge per particle."""
    rng = np.random.default_rng(2)
    d1 = rng.random((3, 100_000))
    d2 = 1.0 - d1
    w = [d2[0] * d2[1] * d2[2], d1[0] * d2[1] * d2[2], d2[0] * d1[1] * d2[2], d1[0] * d1[1] * d2[2],
         d2[0] * d2[1] * d1[2], d1[0] * d2[1] * d1[2], d2[0] * d1[1] * d1[2], d1[0] * d1[1] * d1[2]]
    tot = sum(fix_weight(x).astype(np.int64) for x in w)
    assert np.max(np.abs(tot - (1 << 44))) <= 8
