"""Pins the oracle's integer work (bit-exact) with the reference's own
known-answer tests and closed forms. CPU only."""
import math

import numpy as np
import pytest


def test_factorize_table(oracle):
    # plan_test.go:102-126
    cases = {16: [4, 4], 8: [4, 2], 12: [4, 3], 1296: [4, 4, 3, 3, 3, 3], 1000: [4, 2, 5, 5, 5], 13: [13]}
    for n, want in cases.items():
        got = oracle.factorize(n)
        assert got == want, (n, got)
        assert math.prod(got) == n
    assert oracle.factors_are_small(13 * 13)
    assert not oracle.factors_are_small(17)
    assert not oracle.factors_are_small(34)
    # the schedules of config-2 lengths
    assert oracle.factorize(6720) == [4, 4, 4, 3, 5, 7]
    assert oracle.factorize(1008) == [4, 4, 3, 3, 7]
    assert oracle.factorize(20580) == [4, 3, 5, 7, 7, 7]


def test_primitive_root_and_is_prime(oracle):
    # rader_test.go:130-162
    known = {2: 1, 3: 2, 5: 2, 7: 3, 11: 2, 13: 2, 17: 3, 23: 5}
    for p, g in known.items():
        assert oracle.primitive_root(p) == g
    for p in [17, 31, 101, 257, 1009, 10007]:
        g = oracle.primitive_root(p)
        seen = set()
        x = 1
        for _ in range(p - 1):
            assert x not in seen
            seen.add(x)
            x = x * g % p
    assert oracle.primitive_root(1009) == 11
    assert oracle.primitive_root(10007) == 5
    cases = {0: False, 1: False, 2: True, 3: True, 4: False, 9: False, 17: True, 25: False,
             10007: True, 10006: False, 9973: True}
    for n, want in cases.items():
        assert oracle.is_prime(n) == want
    # brute force agreement up to 2000
    for n in range(2000):
        assert oracle.is_prime(n) == (n > 1 and all(n % d for d in range(2, int(n ** 0.5) + 1)))


def test_routing(oracle):
    # rader_test.go:165-180 and the config-2 routes
    assert oracle.route(641) == "bluestein"
    assert oracle.route(769) == "rader"
    assert oracle.route(4096) == "iterative"
    assert oracle.route(6720) == "mixedradix"
    assert oracle.route(1009) == "rader"
    assert oracle.route(10007) == "rader"
    assert oracle.route(20014) == "bluestein"
    assert oracle.route(1) == "trivial"
    assert oracle.route(0) == "trivial"


def test_conv_length_selection(oracle):
    # radix7_test.go:109-184
    for n in [769, 2017]:
        _, _, cl, cyc = oracle.rader_info(n)
        assert cyc and cl == n - 1
    for n in [5003, 9973, 10007]:
        _, _, cl, cyc = oracle.rader_info(n)
        assert not cyc and cl >= 2 * (n - 1)
        assert oracle.conv_cost(cl)[1]
    # the lengths plan.go's code computes (its comments quote other values)
    assert oracle.rader_info(10007)[2] == 20580
    assert oracle.rader_info(9973)[2] == 20580
    assert oracle.rader_info(5003)[2] == 10290
    assert oracle.rader_info(1009)[2:] == (1008, True)
    for m in [11, 22, 121, 2 * 277, 5003, 10006]:
        assert not oracle.conv_cost(m)[1]
    for m in [1, 2, 4, 8, 7, 10080, 20580, 21952]:
        c, ok = oracle.conv_cost(m)
        assert ok and (m == 1 or c > 0)
    for lo in [1, 2, 16, 10004, 19944, 20012, 100000]:
        g = oracle.best_conv_len(lo)
        gc, ok = oracle.conv_cost(g)
        assert g >= lo and ok
        hi = lo + (lo * 4 + 9) // 10
        for m in range(lo, hi + 1):
            c, ok = oracle.conv_cost(m)
            assert not (ok and c < gc)
    assert oracle.best_conv_len(0) >= 1


def test_next_smooth_conv(oracle):
    def smooth(m):
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        return m == 1
    for lo in [0, 1, 2, 7, 11, 13, 100, 10004, 19944, 20012]:
        g = oracle.next_smooth_conv(lo)
        want = max(lo, 1)
        assert g >= want and smooth(g)
        assert not any(smooth(m) for m in range(want, g))


def test_rader_permutations(oracle):
    for n in [17, 101, 769, 1009, 10007]:
        perm, iperm, cl, cyc = oracle.rader_info(n)
        q = n - 1
        g = oracle.primitive_root(n)
        want = [pow(g, k, n) for k in range(q)]
        assert list(perm) == want
        assert sorted(perm) == list(range(1, n))
        # iperm[q] = g^{-q} mod n
        assert list(iperm) == [perm[(q - k) % q] for k in range(q)]
        for k in range(q):
            assert perm[k] * iperm[k] % n == 1


def test_bluestein_chirp_indices(oracle):
    for n in [17, 641, 20014]:
        m, ph = oracle.bluestein_info(n)
        assert m == 1 << (2 * n - 2).bit_length()
        assert m >= 2 * n - 1 and m // 2 < 2 * n - 1
        j = np.arange(n, dtype=object)
        assert [int(v) for v in ph] == [int(v) for v in (j * j) % (2 * n)]


def test_bit_reverse(oracle):
    # internal/kernels/kernels_test.go:21-31
    x = np.arange(8).astype(np.complex128)
    assert list(oracle.bit_reverse(x).real.astype(int)) == [0, 4, 2, 6, 1, 5, 3, 7]
    for n in [2, 4, 16, 1024]:
        b = n.bit_length() - 1
        want = [int(format(i, "0%db" % b)[::-1], 2) for i in range(n)]
        assert list(oracle.bit_reverse(np.arange(n).astype(np.complex128)).real.astype(int)) == want


def test_digit_reversal(oracle):
    # iterative.go:65-103: odd log2 leads with a radix-2 stage
    assert oracle.it_radices(4096) == [4] * 6
    assert oracle.it_radices(512) == [2, 4, 4, 4, 4]
    for n in [8, 32, 64, 512, 4096]:
        rads = oracle.it_radices(n)
        rev = oracle.bitrev_for_radices(n, rads)
        assert sorted(rev) == list(range(n))
        for i in range(n):
            x, r = i, 0
            for rad in rads:
                r = r * rad + x % rad
                x //= rad
            assert rev[i] == r
    # pure radix-2 digit reversal is the bit reversal
    rev = oracle.bitrev_for_radices(16, [2, 2, 2, 2])
    assert list(rev) == [0, 8, 4, 12, 2, 10, 6, 14, 1, 9, 5, 13, 3, 11, 7, 15]


def test_spectrogram_frame_counts(oracle):
    # freq_test.go:150-177: 20 samples, segment 8, overlap 4 -> 4 frames
    w = oracle.window("hann", 8)
    assert oracle.spectrogram(np.zeros(20), 8, 4, w).shape == (4, 5)
    # the 2^26-sample workload: floor((2^26 - 1024) / 512) + 1 = 131071 frames
    n, seg, ov = 1 << 26, 1024, 512
    assert (n - seg) // (seg - ov) + 1 == 131071
    assert oracle.spectrogram(np.zeros(10), 16, 0, oracle.window("hann", 16)).shape[0] == 0
    with pytest.raises(ValueError, match="segment length must be positive"):
        oracle.spectrogram(np.zeros(4), 0, 0, [])
    with pytest.raises(ValueError, match="window length must equal segment length"):
        oracle.spectrogram(np.zeros(4), 4, 0, [1, 1])
    with pytest.raises(ValueError, match="overlap must satisfy"):
        oracle.spectrogram(np.zeros(4), 4, 4, [1, 1, 1, 1])
