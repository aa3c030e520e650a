"""Seeded generators and host format helpers of libarmsparse (host code, no GPU)."""
import numpy as np
import pytest

import oracle
from armsparse_amd import data

M64 = (1 << 64) - 1
G = 0x9E3779B97F4A7C15


def _mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _rng(seed, stream, i):
    key = _mix((seed + G * (stream + 1)) & M64)
    return _mix((key + G * (i + 1)) & M64)


def test_rng_matches_published_formula():
    from armsparse_amd._lib import lib

    for seed, stream, i in [(0, 0, 0), (1, 2, 3), (0xA5A50001, 16, 123456789), (M64, 7, 2**40)]:
        assert lib.armsp_rng_u64(seed, stream, i) == _rng(seed, stream, i)
    u = data.host_uniform(1000, 42, 16)
    expect = np.array([((_rng(42, 16, i) >> 40) * 2.0**-24) * 2 - 1 for i in range(1000)])
    assert np.array_equal(u.astype(np.float64), expect)  # exact: U(-1,1) needs no rounding


def test_uniform_csr_exact_and_canonical():
    a = data.uniform_csr(300, 500, 4000, seed=5)
    assert a.nnz == 4000 and data.csr_check(a) == 0
    assert a.rowptr[0] == 0 and a.rowptr[-1] == 4000
    assert not np.any(a.values == 0) and np.all(np.abs(a.values) <= 1)
    b = data.uniform_csr(300, 500, 4000, seed=5)
    assert np.array_equal(a.colind, b.colind) and np.array_equal(a.values, b.values)
    c = data.uniform_csr(300, 500, 4000, seed=6)
    assert not np.array_equal(a.colind, c.colind)
    # the structure contract: toCSR of the dense form returns the same arrays
    rp, ci, va = oracle.toCSR(oracle.densify(a.rowptr, a.colind, a.values, 500))
    assert np.array_equal(rp, a.rowptr) and np.array_equal(ci, a.colind) and np.array_equal(va, a.values)


def test_rmat_power_law():
    a = data.rmat_csr(12, 1 << 16, seed=3)
    assert a.nnz == 1 << 16 and data.csr_check(a) == 0
    lens = a.row_lengths()
    # P(row bit = 0) = a + b = 0.76 per level: the first half holds ~76% of the edges
    frac = lens[: len(lens) // 2].sum() / a.nnz
    assert 0.72 < frac < 0.80
    assert lens[0] > 20 * np.median(lens[lens > 0])  # hub row
    assert (lens == 0).mean() > 0.1  # many empty rows


def test_powerlaw_graph_symmetric():
    g = data.powerlaw_graph(2000, 5000, seed=1)
    assert g.nnz == 10000 and data.csr_check(g) == 0
    rows = np.repeat(np.arange(2000), np.diff(g.rowptr))
    assert not np.any(rows == g.colind)
    key = set(zip(rows.tolist(), g.colind.tolist()))
    assert all((c, r) in key for r, c in key)


def test_gcn_graph_matches_proc_convention():
    g = data.gcn_graph(500, 1200, seed=2)
    n = 500
    assert g.rowptr[-1] == 2 * 1200 + n  # directed edges plus self loops
    rows = np.repeat(np.arange(n), np.diff(g.rowptr))
    assert np.all(np.isin(np.arange(n) * (n + 1), rows.astype(np.int64) * n + g.colind))
    # A + I of the self-loop-free generator, built independently with scipy
    import scipy.sparse as sp
    a = data.powerlaw_graph(n, 1200, seed=2)
    ref = (sp.csr_matrix((np.ones(a.nnz), a.colind, a.rowptr), shape=(n, n)) + sp.identity(n, format="csr")).tocsr()
    ref.sort_indices()
    assert np.array_equal(g.rowptr, ref.indptr) and np.array_equal(g.colind, ref.indices)
    assert np.array_equal(g.colptr, g.rowptr) and np.array_equal(g.rowind, g.colind)


@pytest.mark.parametrize("kind,avg,param", [("uniform", 5.5, 0), ("rmat", 8, 0.6), ("banded", 12, 16),
                                            ("blockdiag", 0, 3)])
def test_corpus_generators(kind, avg, param):
    a = data.corpus_csr(kind, 1024, avg, param, seed=9)
    assert data.csr_check(a) == 0
    lens = a.row_lengths()
    rows = np.repeat(np.arange(1024), lens)
    if kind == "banded":
        assert np.all(np.abs(a.colind - rows) <= param)
    if kind == "blockdiag":
        assert np.all(lens % 32 == 0) and np.all(lens >= 32) and np.all(lens <= 32 * param)
        assert np.all(np.isin(rows // 32, a.colind // 32))  # diagonal block present
    if kind == "uniform":
        assert abs(lens.mean() - avg) < 0.5
    if kind == "banded":  # draws inside the band are merged when they collide
        assert 0.75 * avg < lens.mean() <= avg


def test_vcsr_converter_matches_restated_toVectorAwareCSR():
    a = data.uniform_csr(64, 96, 400, seed=4)
    v = data.csr_to_vcsr(a)
    rp, ci, va = oracle.toVectorAwareCSR(oracle.densify(a.rowptr, a.colind, a.values, 96))
    assert np.array_equal(v.rowptr, rp) and np.array_equal(v.colind, ci) and np.array_equal(v.values, va)
    v2 = data.csr_to_vcsr(a, threshold=0.5)
    rp2, ci2, va2 = oracle.toVectorAwareCSR(oracle.densify(a.rowptr, a.colind, a.values, 96), threshold=0.5)
    assert np.array_equal(v2.rowptr, rp2) and np.array_equal(v2.colind, ci2) and np.array_equal(v2.values, va2)


def test_csr_check_codes():
    good = data.CSR(np.array([0, 2, 3], np.int32), np.array([0, 2, 1], np.int32), None, (2, 3))
    assert data.csr_check(good) == 0
    bad = [
        (np.array([1, 2, 3]), np.array([0, 2, 1]), 1),
        (np.array([0, 2, 1]), np.array([0, 2, 1]), 2),
        (np.array([0, 2, 3]), np.array([0, 3, 1]), 4),
        (np.array([0, 2, 3]), np.array([2, 0, 1]), 5),
    ]
    for rp, ci, code in bad:
        assert data.csr_check(data.CSR(rp.astype(np.int32), ci.astype(np.int32), None, (2, 3))) == code
