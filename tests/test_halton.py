"""Halton sampler: the reference's radical inverse (src/util.cpp:22-32) and
the scrambled extension, pinned by closed forms, the stratification
property and an exact rational evaluation; product == oracle bit for bit."""
from fractions import Fraction

import numpy as np
import pytest

from globillum_amd import _lib


def test_radical_inverse_known_values(oracle):
    # closed form: base 2, n = 1..8 and base 3, n = 1..3
    exp2 = [0.5, 0.25, 0.75, 0.125, 0.625, 0.375, 0.875, 0.0625]
    assert [oracle.radical_inverse(n, 2) for n in range(1, 9)] == exp2
    assert oracle.radical_inverse(1, 3) == pytest.approx(1 / 3, rel=1e-7)
    assert oracle.radical_inverse(2, 3) == pytest.approx(2 / 3, rel=1e-7)
    assert oracle.radical_inverse(3, 3) == pytest.approx(1 / 9, rel=1e-7)
    assert oracle.radical_inverse(0, 5) == 0.0


def test_primes(oracle):
    first = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47]
    assert [oracle.prime(d) for d in range(15)] == first
    assert oracle.prime(227) == 1439 and oracle.prime(99) == 541  # 228th, 100th prime


@pytest.mark.parametrize("dim,seed", [(0, 0), (1, 7), (5, 3), (40, 11), (227, 1)])
def test_permutation_is_bijection(oracle, dim, seed):
    p = oracle.halton_perm(dim, seed)
    assert sorted(p.tolist()) == list(range(oracle.prime(dim)))


def test_permutations_depend_on_seed(oracle):
    assert not np.array_equal(oracle.halton_perm(30, 0), oracle.halton_perm(30, 1))
    assert np.array_equal(oracle.halton_perm(30, 5), oracle.halton_perm(30, 5))


def _exact(n, p, perm):
    """sum_k perm[d_k] p^-(k+1) + perm[0] p^-K / (p-1), as a rational."""
    r = Fraction(0)
    k = 0
    while n > 0:
        n, d = divmod(n, p)
        r += Fraction(int(perm[d]), p ** (k + 1))
        k += 1
    return r + Fraction(int(perm[0]), (p - 1) * p ** k)


@pytest.mark.parametrize("dim", [0, 1, 2, 7, 33, 100])
def test_scrambled_matches_exact_rational(oracle, dim):
    rng = np.random.default_rng(dim)
    p = oracle.prime(dim)
    perm = oracle.halton_perm(dim, 9)
    for n in [0, 1, 2, p - 1, p, p * p + 3] + rng.integers(0, 2 ** 32, 50).tolist():
        got = np.float32(oracle.halton_sample(int(n), dim, 9))
        ref = np.float32(float(_exact(int(n), p, perm)))
        assert abs(int(got.view(np.int32)) - int(ref.view(np.int32))) <= 1, (n, got, ref)


@pytest.mark.parametrize("dim", [0, 1, 3, 9])
def test_scrambled_stratification(oracle, dim):
    # the first p^2 points of a scrambled van der Corput sequence are
    # j/p^2 + tail with j = perm[d0]*p + perm[d1] running over all cells and
    # tail = perm[0]/(p-1)/p^2 (the digits beyond the second)
    p = oracle.prime(dim)
    m = p * p
    perm0 = int(oracle.halton_perm(dim, 4)[0])
    v = np.array([oracle.halton_sample(n, dim, 4) for n in range(m)], np.float64)
    cells = np.rint(v * m - perm0 / (p - 1)).astype(int)
    assert np.all(np.abs(v * m - perm0 / (p - 1) - cells) < 1e-4 * m)
    assert sorted(cells.tolist()) == list(range(m))


def test_product_matches_oracle_bitwise(oracle):
    lib = _lib.load()
    rng = np.random.default_rng(1)
    for _ in range(400):
        n = int(rng.integers(0, 2 ** 32))
        dim = int(rng.integers(0, 228))
        seed = int(rng.integers(0, 2 ** 31))
        a = np.float32(lib.gi_halton_sample(n, dim, seed))
        b = np.float32(oracle.halton_sample(n, dim, seed))
        assert a.view(np.uint32) == b.view(np.uint32)
    for n, base in [(1, 2), (77, 3), (12345, 7), (2 ** 31 + 5, 13)]:
        assert np.float32(lib.gi_radical_inverse(n, base)) == np.float32(oracle.radical_inverse(n, base))
    perm = np.zeros(oracle.prime(50), np.uint16)
    assert lib.gi_halton_permutation(50, 17, _lib.ptr(perm)) == 0
    assert np.array_equal(perm, oracle.halton_perm(50, 17))
    assert lib.gi_halton_prime(-1) == -1 and lib.gi_halton_prime(227) == 1439


def test_every_dimension_needs_at_most_seven_chunks():
    # the device splits an index n < 2^32 into at most 7 chunks of m digits,
    # P = p^m <= 1024; that requires P^7 >= 2^32 for every dimension used
    # (4 camera dimensions + 7 per vertex, 32 vertices)
    lib = _lib.load()
    for dim in range(4 + 7 * 32):
        p = lib.gi_halton_prime(dim)
        P = p
        while P * p <= 1024:
            P *= p
        assert P ** 7 >= 2 ** 32, (dim, p, P)


def test_sampler_tables_fit_the_lds_copy():
    # ray generation copies the chunk tables of dimensions 0..3 into LDS, the
    # untextured shading those of its vertex's 7 dimensions: 10240 entries
    # hold every such window of a depth-32 render (the largest, dimensions
    # 221..227, has 9959)
    lib = _lib.load()
    sizes = []
    for dim in range(4 + 7 * 32):
        p = lib.gi_halton_prime(dim)
        P = p
        while P * p <= 1024:
            P *= p
        sizes.append(P)
    windows = [sum(sizes[0:4])] + [sum(sizes[d:d + 7]) for d in range(4, 4 + 7 * 32, 7)]
    assert windows[0] == 2721 and max(windows) == 9959 and max(windows) <= 10240


def test_fixed_chunk_count_keeps_the_rational(oracle):
    # the kernels split every index n < 2^32 into the same kmax chunks of its
    # dimension (the least K with P^K >= 2^32), zero chunks above the leading
    # one included: each zero chunk multiplies A and B by P, so A / B is the
    # digit loop's rational, and A, B stay below 2^53 (exact in FP64) for
    # every dimension of a depth-32 render
    from fractions import Fraction

    lib = _lib.load()
    rng = np.random.default_rng(5)
    seed = 7
    for dim in range(4 + 7 * 32):
        p = lib.gi_halton_prime(dim)
        P, m = p, 1
        while P * p <= 1024:
            P *= p
            m += 1
        kmax, pk = 0, 1
        while pk < 2 ** 32:
            pk *= P
            kmax += 1
        assert pk * p < 2 ** 53, (dim, p, P, kmax)
        perm = oracle.halton_perm(dim, seed).astype(np.int64)
        for n in [0, 1, P - 1, P, P * P, 2 ** 32 - 1] + [int(x) for x in rng.integers(0, 2 ** 32, 4)]:
            rev, pk_d, x = 0, 1, n
            while x:
                rev = rev * p + int(perm[x % p])
                pk_d *= p
                x //= p
            exact = Fraction(rev * (p - 1) + int(perm[0]), pk_d * (p - 1))
            rev, x = 0, n
            for _ in range(kmax):
                c = x % P
                x //= P
                v, y = 0, c
                for _ in range(m):  # the chunk table entry: least significant digit first
                    v = v * p + int(perm[y % p])
                    y //= p
                rev = rev * P + v
            assert x == 0
            A, B = rev * (p - 1) + int(perm[0]), pk * (p - 1)
            assert A < 2 ** 53 and B < 2 ** 53 and Fraction(A, B) == exact, (dim, n)
