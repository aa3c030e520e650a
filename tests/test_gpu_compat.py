"""GPU tests of the reference-name drop-ins (armsparse_compat.h): every shim is
called through its C++ symbol on host arrays, as a reference driver linked
against libarmsparse.so would call it, and compared bit for bit with the
oracle's restatement of the same reference function, on CSR built by both
toCSR (SDMM.cpp:355-370, increasing columns) and toCSR2 (SDMM.cpp:371-394,
odd rows in decreasing column order)."""
import ctypes

import numpy as np
import pytest

import oracle
from armsparse_amd import _lib, data

pytestmark = pytest.mark.gpu

lib = ctypes.CDLL(_lib.LIB_PATH)
_P = ctypes.c_void_p
_I = ctypes.c_int
# Itanium-ABI names of the declarations in armsparse_compat.h
SDMM_SHIMS = {
    "rowBaseSpmm": "_Z11rowBaseSpmmPiS_PfS0_S0_iiii",
    "rowBaseSpmm2": "_Z12rowBaseSpmm2PiS_PfS0_S0_iiii",
    "rowBaseSpmm3": "_Z12rowBaseSpmm3PiS_PfS0_S0_iiii",
    "rowWiseBlocking": "_Z15rowWiseBlockingPiS_PfS0_S0_iiii",
    "rowWiseRegisterBlocking": "_Z23rowWiseRegisterBlockingPiS_PfS0_S0_iiii",
    "irregular": "_Z9irregularPiS_PfS0_S0_iiii",
    "origin": "_Z6originPiS_PfS0_S0_iiii",
}
for _sym in SDMM_SHIMS.values():
    getattr(lib, _sym).argtypes = [_P, _P, _P, _P, _P, _I, _I, _I, _I]
    getattr(lib, _sym).restype = None
_MUL3 = lib._Z18sparse_matrix_mul3PiS_PfS_S_S0_S0_iiii
_MUL3.argtypes = [_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I]
_SDV = lib._Z19sparse_dense_vectorPfS_PiS0_S_ii
_SDV.argtypes = [_P, _P, _P, _P, _P, _I, _I]
_CSR_KERNEL = lib._Z15csr_spmm_kerneliiPiS_PfS0_S0_
_CSR_KERNEL.argtypes = [_I, _I, _P, _P, _P, _P, _P]
_DATAREUSE = lib._Z13DataReuseSpmmPiS_PfS0_S0_iiiiPb
_DATAREUSE.argtypes = [_P, _P, _P, _P, _P, _I, _I, _I, _I, _P]


def p(a):
    return ctypes.c_void_p(a.ctypes.data)


def to_csr2(dense):
    """toCSR2 (SDMM.cpp:371-394): even rows scanned left to right, odd rows right to left."""
    rows, cols = dense.shape
    rp, ci, va = [0], [], []
    for r in range(rows):
        order = range(cols) if r % 2 == 0 else range(cols - 1, -1, -1)
        for c in order:
            if dense[r, c] != 0:
                ci.append(c)
                va.append(dense[r, c])
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(va, np.float32)


def sparse_dense(rows, cols, density, seed):
    """A dense matrix with the reference generators' value set ((rand%20000-10000)/1000, SDMM.cpp:329-342)."""
    r = np.random.default_rng(seed)
    d = ((r.integers(0, 20000, (rows, cols)) - 10000) / 1000).astype(np.float32)
    d[r.random((rows, cols)) >= density] = 0.0
    return d


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("builder", ["toCSR", "toCSR2"])
@pytest.mark.parametrize("name", list(SDMM_SHIMS))
def test_sdmm_shim_bitexact(name, builder):
    """C (+)= A·B through the shim on host arrays equals the restated reference kernel bit for bit."""
    M, K, N = 96, 80, 40
    dense = sparse_dense(M, K, 0.15, 7)
    rp, ci, va = oracle.toCSR(dense) if builder == "toCSR" else to_csr2(dense)
    B = np.random.default_rng(8).uniform(-1, 1, (K, N)).astype(np.float32)
    C0 = np.random.default_rng(9).uniform(-1, 1, (M, N)).astype(np.float32)
    C = C0.copy()
    getattr(lib, SDMM_SHIMS[name])(p(rp), p(ci), p(va), p(B), p(C), M, K, N, 8)
    expect = getattr(oracle, name if name != "rowBaseSpmm2" else "rowBaseSpmm3")(rp, ci, va, B, C0.copy())
    assert np.array_equal(bits(C), bits(expect)), name
    if name not in ("origin",):
        assert not np.array_equal(C, C0)


@pytest.mark.parametrize("builder", ["toCSR", "toCSR2"])
def test_sparse_matrix_mul3_shim_bitexact(builder):
    M, K, N = 64, 48, 72
    da, db = sparse_dense(M, K, 0.2, 10), sparse_dense(K, N, 0.2, 11)
    a = oracle.toCSR(da) if builder == "toCSR" else to_csr2(da)
    b = oracle.toCSR(db) if builder == "toCSR" else to_csr2(db)
    C0 = np.random.default_rng(12).uniform(-1, 1, (M, N)).astype(np.float32)
    C = C0.copy()
    _MUL3(p(a[0]), p(a[1]), p(a[2]), p(b[0]), p(b[1]), p(b[2]), p(C), M, K, N, 8)
    expect = oracle.sparse_matrix_mul3(a, b, C0.copy())
    assert np.array_equal(bits(C), bits(expect))


def test_sparse_dense_vector_shim_bitexact():
    A, _, x = data.spmm_small()
    v = data.csr_to_vcsr(A)
    y0 = np.random.default_rng(13).uniform(-1, 1, A.shape[0]).astype(np.float32)
    y = y0.copy()
    xs = np.ascontiguousarray(x[: v.shape[1]])
    _SDV(p(xs), p(y), p(v.rowptr), p(v.colind), p(v.values), A.shape[0], 8)
    expect = oracle.sparse_dense_vector(xs, y0.copy(), v.rowptr, v.colind, v.values)
    assert np.array_equal(bits(y), bits(expect))


@pytest.mark.parametrize("builder", ["toCSR", "toCSR2"])
def test_csr_spmm_kernel_shim_bitexact(builder):
    """csr_spmm_kernel (spmm.cpp:49-71) on a zeroed output, as the drop-in op calls it."""
    M, K, N = 70, 70, 24
    dense = sparse_dense(M, K, 0.1, 14)
    rp, ci, va = oracle.toCSR(dense) if builder == "toCSR" else to_csr2(dense)
    B = np.random.default_rng(15).uniform(-1, 1, (K, N)).astype(np.float32)
    C = np.zeros((M, N), np.float32)
    _CSR_KERNEL(M, N, p(rp), p(ci), p(va), p(B), p(C))
    expect = oracle.rowBaseSpmm3(rp, ci, va, B, np.zeros((M, N), np.float32))
    assert np.array_equal(bits(C), bits(expect))


def test_datareuse_shim_bitexact():
    M, K, N = 40, 30, 16
    dense = sparse_dense(M, K, 0.3, 16)
    rp, ci, va, fl = oracle.toDataReuseFormat(dense)
    B = np.random.default_rng(17).uniform(-1, 1, (K, N)).astype(np.float32)
    C0 = np.random.default_rng(18).uniform(-1, 1, (M, N)).astype(np.float32)
    C = C0.copy()
    flb = np.ascontiguousarray(fl.astype(np.bool_))
    _DATAREUSE(p(rp), p(ci), p(va), p(B), p(C), M, K, N, 8, p(flb))
    expect = oracle.DataReuseSpmm(rp, ci, va, fl, B, C0.copy())
    assert np.array_equal(bits(C), bits(expect))
