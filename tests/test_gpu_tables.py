"""GPU: the product's integer index work, bit for bit against the oracle.

A different primitive root, a shifted inverse permutation or a wrong chirp
index can still give a correct DFT (or an error the floating-point parity
bars do not see), so these tables are read back from the plan and compared
exactly: Rader's generator and its permutations perm[k] = g^k mod n and
iperm[k] = perm[(q - k) mod q] (rader.go:78-86), the device copies the kernels
index with; Bluestein's chirp phase indices j^2 mod 2n (bluestein.go:35-41),
the ones the plan's chirp tables were built from.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RADER = [769, 1009, 2017, 5003, 10007, 20011, 600011]
BLUE = [641, 20014, 600014, 2 * 1009]


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    from fft_amd import _lib as L
    return D, L, torch


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("n", RADER)
def test_rader_tables_bit_exact(dev, oracle, n, prec):
    D, L, torch = dev
    p = D.C2CPlan(n, 2, torch.complex128 if prec == "c128" else torch.complex64)
    perm, iperm, cl_ref, cyclic_ref = oracle.rader_info(n)
    assert p.index_table(n, L.TABLE_RADER_GENERATOR).tolist() == [oracle.primitive_root(n)]
    got_perm = p.index_table(n, L.TABLE_RADER_PERM)
    got_iperm = p.index_table(n, L.TABLE_RADER_IPERM)
    assert got_perm.dtype == np.int64 and np.array_equal(got_perm, perm)
    assert np.array_equal(got_iperm, iperm)
    # the convolution: cyclic exactly when the reference's is (q 13-smooth);
    # its length is the product's own choice, at least 2q when linear
    cl, cyclic = p.index_table(n, L.TABLE_RADER_CONV).tolist()
    q = n - 1
    assert bool(cyclic) == cyclic_ref
    assert (cl == q) if cyclic else (cl >= 2 * q)


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("n", BLUE)
def test_bluestein_chirp_index_bit_exact(dev, oracle, n, prec):
    D, L, torch = dev
    p = D.C2CPlan(n, 2, torch.complex128 if prec == "c128" else torch.complex64)
    m_ref, ph = oracle.bluestein_info(n)
    got = p.index_table(n, L.TABLE_CHIRP_INDEX)
    assert np.array_equal(got, ph)
    # closed form, independently of the oracle: j^2 mod 2n in exact integers
    j = np.arange(n, dtype=object)
    assert got.tolist() == [int(v) for v in (j * j) % (2 * n)]
    (m,) = p.index_table(n, L.TABLE_BLUESTEIN_CONV).tolist()
    if prec == "c128":
        assert m == m_ref  # the reference's power of two (reference-twiddle mode)
    else:
        x = m
        for f in (2, 3, 5, 7):
            while x % f == 0:
                x //= f
        assert x == 1 and m >= 2 * n - 1


def test_tables_inside_mixed_and_nd_plans(dev, oracle):
    # the same tables through the mixed-length plan's groups and an N-D axis
    D, L, torch = dev
    mp = D.ManyPlan([1024, 1009, 20014, 10007, 1009], torch.complex64)
    for n in (1009, 10007):
        assert np.array_equal(mp.index_table(n, L.TABLE_RADER_PERM), oracle.rader_info(n)[0])
        assert np.array_equal(mp.index_table(n, L.TABLE_RADER_IPERM), oracle.rader_info(n)[1])
    assert np.array_equal(mp.index_table(20014, L.TABLE_CHIRP_INDEX), oracle.bluestein_info(20014)[1])
    nd = D.NDPlan((16, 1009), L.ND_C2C, 1, torch.complex128)
    assert np.array_equal(nd.index_table(1009, L.TABLE_RADER_PERM), oracle.rader_info(1009)[0])
    with pytest.raises(L.GofftError):
        mp.index_table(4096, L.TABLE_RADER_PERM)  # no such length
    with pytest.raises(L.GofftError):
        mp.index_table(1024, L.TABLE_RADER_PERM)  # not routed to Rader
