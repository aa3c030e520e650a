"""The 2-row DataReuse format (toDataReuseFormat, suitesparse_sdmm.cpp:150-179)
and its SpMM (DataReuseSpmm, :25-43): the oracle's restatement against the
golden 8×7 fixture and an FP64 product, and the library's CSR-side converter
against the restatement on dense matrices."""
import numpy as np
import pytest

import oracle
from armsparse_amd import data
from conftest import FIX8_COLIND, FIX8_ROWPTR, FIX8_ROWSUMS, FIX8_VAL, random_csr


def _dense(rp, ci, va, k):
    return oracle.densify(rp, ci, va, k)


def test_golden_fixture_row_sums():
    """SDMM.cpp:231-247's 8×7 matrix with B = 1: the DataReuse product gives the
    reference's row sums."""
    rp, ci, va, fl = oracle.toDataReuseFormat(_dense(FIX8_ROWPTR, FIX8_COLIND, FIX8_VAL, 7))
    C = np.zeros((8, 6), np.float32)
    oracle.DataReuseSpmm(rp, ci, va, fl, np.ones((7, 6), np.float32), C)
    assert np.array_equal(C, np.repeat(FIX8_ROWSUMS[:, None], 6, axis=1))


@pytest.mark.parametrize("m,k,density,seed", [(8, 7, None, 0), (301, 200, 0.05, 1), (64, 64, 0.5, 2),
                                              (7, 9, 1.0, 3), (100, 50, 0.0, 4)])
def test_converter_equals_the_dense_restatement(m, k, density, seed):
    if density is None:
        rp, ci, va = FIX8_ROWPTR, FIX8_COLIND, FIX8_VAL
    else:
        rp, ci, va = random_csr(m, k, density, seed, empty_rows=0.1)
    dense = _dense(rp, ci, va, k)
    want = oracle.toDataReuseFormat(dense)
    rp, ci, va = oracle.toCSR(dense)  # the fixture holds an explicit 0; a dense matrix cannot
    got = data.csr_to_datareuse(data.CSR(rp, ci, va, (m, k)))
    for w, g in zip(want, got):
        assert np.array_equal(w, g)
    # flagged entries sit in even rows, right behind an entry of the same column
    frp, fci, _, ffl = got
    rows = np.repeat(np.arange(m), np.diff(frp))
    f = np.nonzero(ffl)[0]
    assert np.all(rows[f] % 2 == 0) and np.all(rows[f] + 1 < m)
    assert np.all(ffl[f - 1] == 0) and np.all(fci[f - 1] == fci[f])


@pytest.mark.parametrize("m,k,n,density,seed", [(301, 200, 24, 0.05, 5), (64, 64, 33, 0.5, 6), (9, 9, 4, 1.0, 7)])
def test_restated_product_matches_fp64(m, k, n, density, seed):
    rp, ci, va = random_csr(m, k, density, seed)
    B = np.random.default_rng(seed).standard_normal((k, n)).astype(np.float32)
    frp, fci, fva, ffl = oracle.toDataReuseFormat(_dense(rp, ci, va, k))
    C = np.zeros((m, n), np.float32)
    oracle.DataReuseSpmm(frp, fci, fva, ffl, B, C)
    ref = _dense(rp, ci, va, k).astype(np.float64) @ B.astype(np.float64)
    bound = np.abs(_dense(rp, ci, va, k)).astype(np.float64) @ np.abs(B).astype(np.float64)
    assert np.all(np.abs(C - ref) <= 1e-5 * bound + 1e-30)
