import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


@__import__("functools").lru_cache(maxsize=None)
def rmat_1m_cached():
    """The 1M R-MAT workload, generated once per test session."""
    from armsparse_amd import data

    return data.rmat_1m()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


# The 8×7 CSR fixture written (commented out) in the reference driver,
# SDMM.cpp:231-247: values 0..26, B = all ones, N = 6.
FIX8_ROWPTR = np.array([0, 4, 7, 9, 14, 15, 21, 25, 27], dtype=np.int32)
FIX8_COLIND = np.array([1, 3, 5, 6, 1, 2, 3, 3, 4, 0, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 6, 0, 2, 4, 5, 2, 6],
                       dtype=np.int32)
FIX8_VAL = np.arange(27, dtype=np.float32)
FIX8_ROWSUMS = np.array([6, 15, 15, 55, 14, 105, 90, 51], dtype=np.float32)
# rowBaseSpmm keeps only whole groups of 4 nonzeros per row (SDMM.cpp:33)
FIX8_TRUNC = np.array([6, 0, 0, 42, 0, 66, 90, 0], dtype=np.float32)


def read_mtx(name):
    """MatrixMarket coordinate file → (rowptr, colind, val, (M, K)) canonical CSR (duplicates summed)."""
    import scipy.io
    import scipy.sparse as sp

    m = sp.csr_matrix(scipy.io.mmread(os.path.join(GOLDEN, name)))
    m.sum_duplicates()
    m.sort_indices()
    val = m.data.astype(np.float32) if m.data.dtype != bool else np.ones(m.nnz, np.float32)
    return m.indptr.astype(np.int32), m.indices.astype(np.int32), val, m.shape


@pytest.fixture(scope="session")
def fix8():
    return FIX8_ROWPTR, FIX8_COLIND, FIX8_VAL


def rng(seed):
    return np.random.default_rng(seed)


def random_csr(M, K, density, seed, empty_rows=0.0):
    """Seeded canonical CSR via numpy (test-only helper)."""
    r = rng(seed)
    mask = r.random((M, K)) < density
    if empty_rows:
        mask[r.random(M) < empty_rows] = False
    rows, cols = np.nonzero(mask)
    rp = np.zeros(M + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    val = r.uniform(-1, 1, size=rows.size).astype(np.float32)
    val[val == 0] = 0.5
    return rp, cols.astype(np.int32), val
