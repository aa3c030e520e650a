"""Pins the oracle (CPU restatement of the reference kernels) before any GPU
result is checked against it: golden vectors from the reference source, the
reference's own sparse_matrix_mul3 compiled from its tree (oracle/_ref),
FP64 closed forms, library routines (scipy) and brute force on tiny inputs."""
import numpy as np
import pytest

import oracle
from conftest import FIX8_ROWSUMS, FIX8_TRUNC, random_csr, read_mtx

# ------------------------------------------------------------ golden vectors


def test_fix8_full_row_sums(fix8):
    """SDMM.cpp:231-247 with B = ones: every column of C holds the row sums."""
    rp, ci, va = fix8
    B = np.ones((7, 6), np.float32)
    for kern in (oracle.rowBaseSpmm3, oracle.origin):
        C = np.zeros((8, 6), np.float32)
        kern(rp, ci, va, B, C)
        assert np.array_equal(C, np.repeat(FIX8_ROWSUMS[:, None], 6, axis=1))
    C = oracle.csr_spmm(rp, ci, va, B)
    assert np.array_equal(C[:, 0], FIX8_ROWSUMS)


def test_fix8_blocking_variants(fix8):
    """rowWiseBlocking (M%8, N%8) and rowWiseRegisterBlocking (N%4) on N = 8."""
    rp, ci, va = fix8
    B = np.ones((7, 8), np.float32)
    for kern in (oracle.rowWiseBlocking, oracle.rowWiseRegisterBlocking):
        C = np.zeros((8, 8), np.float32)
        kern(rp, ci, va, B, C)
        assert np.array_equal(C, np.repeat(FIX8_ROWSUMS[:, None], 8, axis=1))


def test_fix8_truncated(fix8):
    """rowBaseSpmm drops the nnz%4 tail of each row and the N%4 columns (SDMM.cpp:33,36)."""
    rp, ci, va = fix8
    C = np.zeros((8, 6), np.float32)
    oracle.rowBaseSpmm(rp, ci, va, np.ones((7, 6), np.float32), C)
    assert np.array_equal(C[:, :4], np.repeat(FIX8_TRUNC[:, None], 4, axis=1))
    assert np.array_equal(C[:, 4:], np.zeros((8, 2), np.float32))


def test_fix8_irregular_only_full_32_blocks(fix8):
    rp, ci, va = fix8
    C = np.zeros((8, 6), np.float32)
    oracle.irregular(rp, ci, va, np.ones((7, 6), np.float32), C)
    assert not C.any()  # N/32 == 0 column blocks
    C = np.zeros((8, 40), np.float32)
    oracle.irregular(rp, ci, va, np.ones((7, 40), np.float32), C)
    assert np.array_equal(C[:, :32], np.repeat(FIX8_ROWSUMS[:, None], 32, axis=1))
    assert not C[:, 32:].any()


# ------------------------------------------------ the reference's own code


@pytest.mark.skipif(oracle.ref is None, reason="oracle/_ref not built (reference tree absent)")
def test_restated_spgemm_matches_reference_build():
    """Restated sparse_matrix_mul3 is bit-exact with the reference's own source
    (numaAware/numaAwareSpMM.cpp, built with FMA contraction as on aarch64)."""
    a = random_csr(300, 200, 0.03, 1)
    b = random_csr(200, 250, 0.03, 2)
    C0 = np.random.default_rng(3).uniform(-1, 1, (300, 250)).astype(np.float32)
    c_ref = oracle.ref_sparse_matrix_mul3(a, b, C0.copy(), threads=4)
    c_orc = oracle.sparse_matrix_mul3(a, b, C0.copy(), threads=4)
    assert np.array_equal(c_ref.view(np.uint32), c_orc.view(np.uint32))
    # and a plausible mistake (one product dropped) is caught
    v = b[2].copy()
    v[0] = 0.0
    c_bad = oracle.sparse_matrix_mul3(a, (b[0], b[1], v), C0.copy(), threads=4)
    assert not np.array_equal(c_ref, c_bad)


# --------------------------------------------------------- closed forms


@pytest.mark.parametrize("N", [1, 3, 16, 64])
def test_spmm_restatement_vs_fp64(N):
    rp, ci, va = random_csr(257, 190, 0.05, 10 + N, empty_rows=0.1)
    r = np.random.default_rng(N)
    B = r.uniform(-1, 1, (190, N)).astype(np.float32)
    C0 = r.uniform(-1, 1, (257, N)).astype(np.float32)
    C = oracle.rowBaseSpmm3(rp, ci, va, B, C0.copy())
    ref = oracle.spmm_f64(rp, ci, va, B, C0)
    dense = oracle.densify(rp, ci, va, 190).astype(np.float64) @ B.astype(np.float64) + C0
    assert np.allclose(ref, dense, rtol=1e-12, atol=1e-12)
    assert np.max(np.abs(C - ref)) < 1e-5
    # C = A·B overwrite form
    Co = oracle.origin(rp, ci, va, B, np.full((257, N), 7.0, np.float32))
    assert np.max(np.abs(Co - oracle.spmm_f64(rp, ci, va, B))) < 1e-5


def test_spmm_special_cases():
    n = 64
    eye = (np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n, np.float32))
    B = np.random.default_rng(0).uniform(-1, 1, (n, 8)).astype(np.float32)
    C0 = np.random.default_rng(1).uniform(-1, 1, (n, 8)).astype(np.float32)
    assert np.array_equal(oracle.rowBaseSpmm3(*eye, B, C0.copy()), C0 + B)  # identity: C += B exactly
    # single nonzero per row: C = round(a*b) exactly (fma(a, b, 0))
    a = np.random.default_rng(2).uniform(-1, 1, n).astype(np.float32)
    C = oracle.rowBaseSpmm3(eye[0], eye[1], a, B, np.zeros((n, 8), np.float32))
    assert np.array_equal(C, (a[:, None] * B).astype(np.float32))
    # empty rows leave C untouched
    rp = np.zeros(n + 1, np.int32)
    C = oracle.rowBaseSpmm3(rp, np.zeros(0, np.int32), np.zeros(0, np.float32), B, C0.copy())
    assert np.array_equal(C, C0)


def test_fp64_check_catches_dropped_term():
    """Planted mistake: the oracle with one nonzero dropped fails the FP64 gate."""
    rp, ci, va = random_csr(100, 100, 0.1, 5)
    B = np.random.default_rng(5).uniform(-1, 1, (100, 16)).astype(np.float32)
    ref = oracle.spmm_f64(rp, ci, va, B)
    bad = va.copy()
    bad[len(bad) // 2] = 0.0
    C = oracle.rowBaseSpmm3(rp, ci, bad, B, np.zeros((100, 16), np.float32))
    assert np.max(oracle.row_rel_err(C, ref)) > 1e-3


def test_numa_sdmm_ksplit():
    """numaAwareSDMM's column split + sumUp equals A·B up to fp32 rounding."""
    rp, ci, va = random_csr(64, 128, 0.1, 7)
    B = np.random.default_rng(7).uniform(-1, 1, (128, 32)).astype(np.float32)
    C = oracle.numa_sdmm(rp, ci, va, B, nodes=4)
    assert np.max(oracle.row_rel_err(C, oracle.spmm_f64(rp, ci, va, B))) < 1e-5


def test_spgemm_fp64_and_dense():
    a = random_csr(60, 50, 0.1, 11)
    b = random_csr(50, 70, 0.1, 12)
    C0 = np.zeros((60, 70), np.float32)
    ref = oracle.spgemm_f64(a, b, C0)
    dense = oracle.densify(*a, 50).astype(np.float64) @ oracle.densify(*b, 70).astype(np.float64)
    assert np.allclose(ref, dense, atol=1e-12)
    C = oracle.sparse_matrix_mul3(a, b, C0.copy())
    assert np.max(np.abs(C - dense)) < 1e-5


# ------------------------------------------------------- SpMV (VCSR)


def _vcsr_brute(x, y, rp, ci, va, fused):
    y = y.copy()
    f32 = np.float32
    for m in range(len(rp) - 1):
        s = [f32(0)] * 4
        for idx in range(rp[m], rp[m + 1], 4):
            k = ci[idx]
            for l in range(4):
                if fused:
                    s[l] = f32(np.float64(va[idx + l]) * np.float64(x[k + l]) + np.float64(s[l]))
                else:
                    s[l] = f32(s[l] + f32(va[idx + l] * x[k + l]))
        y[m] = f32(y[m] + f32(f32(s[2] + s[0]) + f32(s[3] + s[1])))
    return y


@pytest.mark.parametrize("fused", [False, True])
def test_vcsr_restatement_bruteforce(fused):
    dense = np.zeros((40, 64), np.float32)
    r = np.random.default_rng(3)
    mask = r.random(dense.shape) < 0.08
    dense[mask] = r.uniform(-1, 1, mask.sum())
    rp, ci, va = oracle.toVectorAwareCSR(dense)
    assert len(ci) % 4 == 0 and np.all(ci[::4] % 4 == 0)
    x = r.uniform(-1, 1, 64).astype(np.float32)
    y0 = r.uniform(-1, 1, 40).astype(np.float32)
    y = oracle.sparse_dense_vector(x, y0.copy(), rp, ci, va, fused=fused)
    assert np.array_equal(y, _vcsr_brute(x, y0, rp, ci, va, fused))
    assert np.max(np.abs(y - (dense.astype(np.float64) @ x + y0))) < 1e-5


def test_vcsr_threshold_variant():
    """numaAwareSpMV's |v| > 0.1 trigger keeps blocks only for large entries."""
    dense = np.zeros((2, 8), np.float32)
    dense[0, 1] = 0.05  # below threshold: block not stored
    dense[1, 5] = 0.5
    dense[1, 6] = 0.01  # stored as part of a kept block
    rp, ci, va = oracle.toVectorAwareCSR(dense, threshold=0.1)
    assert list(rp) == [0, 0, 4] and list(ci) == [4, 5, 6, 7]
    assert np.array_equal(va, np.array([0, 0.5, 0.01, 0], np.float32))


# ------------------------------------------------------------ structure


def test_toCSR_toCSC_match_scipy():
    import scipy.sparse as sp

    r = np.random.default_rng(9)
    dense = np.where(r.random((50, 37)) < 0.1, r.uniform(-1, 1, (50, 37)), 0).astype(np.float32)
    rp, ci, va = oracle.toCSR(dense)
    s = sp.csr_matrix(dense)
    assert np.array_equal(rp, s.indptr) and np.array_equal(ci, s.indices) and np.array_equal(va, s.data)
    cp, ri, vc = oracle.toCSC(dense)
    t = sp.csc_matrix(dense)
    assert np.array_equal(cp, t.indptr) and np.array_equal(ri, t.indices) and np.array_equal(vc, t.data)
    # stable counting-sort transpose of the CSR equals the dense-scan CSC; twice is the identity
    cp2, ri2, vc2 = oracle.csr_to_csc(rp, ci, va, 37)
    assert np.array_equal(cp2, cp) and np.array_equal(ri2, ri) and np.array_equal(vc2, vc)
    rp3, ci3, va3 = oracle.csr_to_csc(cp2, ri2, vc2, 50)
    assert np.array_equal(rp3, rp) and np.array_equal(ci3, ci) and np.array_equal(va3, va)


def test_sddmm_fp64_vs_numpy():
    rp, ci, _ = random_csr(40, 30, 0.2, 4)
    r = np.random.default_rng(4)
    U = r.standard_normal((40, 12)).astype(np.float32)
    V = r.standard_normal((30, 12)).astype(np.float32)
    s = r.uniform(-1, 1, len(ci)).astype(np.float32)
    rows = np.repeat(np.arange(40), np.diff(rp))
    expect = np.einsum("ek,ek->e", U[rows].astype(np.float64), V[ci].astype(np.float64))
    assert np.allclose(oracle.sddmm_f64(rp, ci, U, V), expect, rtol=1e-12)
    assert np.allclose(oracle.sddmm_f64(rp, ci, U, V, scale=s), expect * s, rtol=1e-12)


@pytest.mark.parametrize("name", ["cage3.mtx", "cage4.mtx", "ibm32.mtx"])
def test_mtx_fixtures(name):
    """Reference fixture matrices (datasets/*.mtx) through restatement vs FP64."""
    rp, ci, va, (M, K) = read_mtx(name)
    B = np.random.default_rng(M).uniform(-1, 1, (K, 8)).astype(np.float32)
    C = oracle.csr_spmm(rp, ci, va, B)
    assert np.max(np.abs(C - oracle.spmm_f64(rp, ci, va, B))) < 1e-5
