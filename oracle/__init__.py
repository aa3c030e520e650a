"""The oracle: CPU restatement of the ARMSparse hot-path kernels (oracle.c)
plus FP64 closed forms, and the reference's own sparse_matrix_mul3 compiled
from the reference tree (oracle/_ref, when it was built).

TEST INFRASTRUCTURE ONLY — imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product package never
imports it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_float, c_int, c_void_p

import numpy as np

# torch bundles its own OpenMP runtime under the soname libgomp.so.1; loading
# torch first makes liboracle.so's (and _ref's) libgomp dependency resolve to
# that same runtime, so the process never holds two OpenMP runtimes.
import torch  # noqa: F401  (imported for its libgomp)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "libref.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")

lib = ctypes.CDLL(LIB_PATH)
_P, _I = c_void_p, c_int
for _n, _a in {
    "orc_num_vector": [_P, _P, _P, _I],
    "orc_rowBaseSpmm3": [_P, _P, _P, _P, _P, _I, _I, _I, _I],
    "orc_rowBaseSpmm": [_P, _P, _P, _P, _P, _I, _I, _I, _I],
    "orc_rowWiseBlocking": [_P, _P, _P, _P, _P, _I, _I, _I, _I],
    "orc_rowWiseRegisterBlocking": [_P, _P, _P, _P, _P, _I, _I, _I, _I],
    "orc_irregular": [_P, _P, _P, _P, _P, _I, _I, _I, _I],
    "orc_origin": [_P, _P, _P, _P, _P, _I, _I, _I, _I],
    "orc_sparse_matrix_mul3": [_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I],
    "orc_sparse_dense_vector": [_P, _P, _P, _P, _P, _I, _I, _I],
    "orc_toCSR": [_P, _I, _I, _P, _P, _P],
    "orc_toCSC": [_P, _I, _I, _P, _P, _P],
    "orc_toDataReuseFormat": [_P, _I, _I, _P, _P, _P, _P],
    "orc_DataReuseSpmm": [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I],
    "orc_toVectorAwareCSR": [_P, _I, _I, _P, _P, _P, c_float],
    "orc_numa_sdmm": [_P, _P, _P, _P, _P, _P, _I, _I, _I, _I],
    "orc_spmm_f64": [_P, _P, _P, _P, _I, _P, _P, _I, _I, _I],
    "orc_spmv_f64": [_P, _P, _P, _P, _P, _P, _I, _I],
    "orc_sddmm_f64": [_P, _P, _P, _P, _I, _P, _I, _I, _P, _I, _I],
    "orc_csr_to_csc": [_P, _P, _P, _I, _I, _P, _P, _P],
    "orc_spgemm_f64": [_P, _P, _P, _P, _P, _P, _P, _P, _I, _I],
}.items():
    getattr(lib, _n).argtypes = _a
    getattr(lib, _n).restype = c_int if _n.startswith("orc_to") else None

ref = None
if os.path.exists(REF_PATH):
    ref = ctypes.CDLL(REF_PATH)
    ref.ref_sparse_matrix_mul3.argtypes = [_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I]
    ref.ref_sparse_matrix_mul3.restype = None

THREADS = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return c_void_p(a.ctypes.data)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


# ------------------------------------------------------ reference kernels
def _spmm_like(name, rowptr, colind, val, B, C, threads):
    rowptr, colind, val, B = _i32(rowptr), _i32(colind), _f32(val), _f32(B)
    assert C.dtype == np.float32 and C.flags["C_CONTIGUOUS"]
    M, K, N = len(rowptr) - 1, B.shape[0], B.shape[1]
    getattr(lib, name)(_p(rowptr), _p(colind), _p(val), _p(B), _p(C), M, K, N, threads or THREADS)
    return C


def rowBaseSpmm3(rowptr, colind, val, B, C, threads=None):
    """C += A·B in place (SDMM.cpp:90-114)."""
    return _spmm_like("orc_rowBaseSpmm3", rowptr, colind, val, B, C, threads)


def rowBaseSpmm(rowptr, colind, val, B, C, threads=None):
    """Truncated C += A·B in place (SDMM.cpp:21-64)."""
    return _spmm_like("orc_rowBaseSpmm", rowptr, colind, val, B, C, threads)


def rowWiseBlocking(rowptr, colind, val, B, C, threads=None):
    return _spmm_like("orc_rowWiseBlocking", rowptr, colind, val, B, C, threads)


def rowWiseRegisterBlocking(rowptr, colind, val, B, C, threads=None):
    return _spmm_like("orc_rowWiseRegisterBlocking", rowptr, colind, val, B, C, threads)


def irregular(rowptr, colind, val, B, C, threads=None):
    return _spmm_like("orc_irregular", rowptr, colind, val, B, C, threads)


def origin(rowptr, colind, val, B, C, threads=None):
    """C = A·B in place (SDMM.cpp:422-448)."""
    return _spmm_like("orc_origin", rowptr, colind, val, B, C, threads)


def csr_spmm(rowptr, colind, val, B, threads=None):
    """The GCN op (gcn/pytorch-custom/spmm.cpp:31-71) on a zeroed output."""
    C = np.zeros((len(rowptr) - 1, B.shape[1]), dtype=np.float32)
    return rowBaseSpmm3(rowptr, colind, val, B, C, threads)


def sparse_matrix_mul3(a, b, C, threads=None):
    """Dense C += A·B in place, A and B CSR (SpMM.cpp:309-327)."""
    (r1, c1, v1), (r2, c2, v2) = [(_i32(x), _i32(y), _f32(z)) for x, y, z in (a, b)]
    M, K, N = len(r1) - 1, len(r2) - 1, C.shape[1]
    lib.orc_sparse_matrix_mul3(_p(r1), _p(c1), _p(v1), _p(r2), _p(c2), _p(v2), _p(C), M, K, N, threads or THREADS)
    return C


def ref_sparse_matrix_mul3(a, b, C, threads=1):
    """The reference's own sparse_matrix_mul3 (numaAware/numaAwareSpMM.cpp), if oracle/_ref was built."""
    if ref is None:
        raise RuntimeError("oracle/_ref/libref.so not built")
    (r1, c1, v1), (r2, c2, v2) = [(_i32(x), _i32(y), _f32(z)) for x, y, z in (a, b)]
    M, K, N = len(r1) - 1, len(r2) - 1, C.shape[1]
    ref.ref_sparse_matrix_mul3(_p(r1), _p(c1), _p(v1), _p(r2), _p(c2), _p(v2), _p(C), M, K, N, threads)
    return C


def sparse_dense_vector(x, y, rowptr, colind, val, fused=False, threads=None):
    """y += A·x in place on the blocked CSR (SpMV.cpp:155-179)."""
    rowptr, colind, val, x = _i32(rowptr), _i32(colind), _f32(val), _f32(x)
    lib.orc_sparse_dense_vector(_p(x), _p(y), _p(rowptr), _p(colind), _p(val), len(rowptr) - 1, threads or THREADS,
                                1 if fused else 0)
    return y


def numa_sdmm(rowptr, colind, val, B, nodes=4):
    """K-split + sumUp of numaAware/numaAwareSDMM.cpp; returns C = A·B."""
    rowptr, colind, val, B = _i32(rowptr), _i32(colind), _f32(val), _f32(B)
    M, K, N = len(rowptr) - 1, B.shape[0], B.shape[1]
    C = np.empty((M, N), dtype=np.float32)
    work = np.empty(((nodes - 1) * M * N,), dtype=np.float32)
    lib.orc_numa_sdmm(_p(rowptr), _p(colind), _p(val), _p(B), _p(C), _p(work), M, K, N, nodes)
    return C


# ------------------------------------------------------------ formats
def toCSR(dense):
    dense = _f32(dense)
    rows, cols = dense.shape
    rp = np.empty(rows + 1, dtype=np.int32)
    nnz = lib.orc_toCSR(_p(dense), rows, cols, _p(rp), None, None)
    ci, va = np.empty(nnz, dtype=np.int32), np.empty(nnz, dtype=np.float32)
    lib.orc_toCSR(_p(dense), rows, cols, _p(rp), _p(ci), _p(va))
    return rp, ci, va


def toDataReuseFormat(dense):
    """(rowptr, colind, val, flag uint8) of toDataReuseFormat on a copy of dense."""
    d = _f32(dense).copy()
    rows, cols = d.shape
    n = int(np.count_nonzero(d))
    rp = np.empty(rows + 1, dtype=np.int32)
    ci, va = np.empty(n, dtype=np.int32), np.empty(n, dtype=np.float32)
    fl = np.empty(n, dtype=np.uint8)
    got = lib.orc_toDataReuseFormat(_p(d), rows, cols, _p(rp), _p(ci), _p(va), _p(fl))
    assert got == n
    return rp, ci, va, fl


def DataReuseSpmm(rowptr, colind, val, flag, B, C, threads=None):
    """C (in place, returned) += A·B through DataReuseSpmm's pair order."""
    B = _f32(B)
    M, N, K = len(rowptr) - 1, B.shape[1], B.shape[0]
    assert C.dtype == np.float32 and C.flags.c_contiguous and C.shape == (M, N)
    lib.orc_DataReuseSpmm(_p(_i32(rowptr)), _p(_i32(colind)), _p(_f32(val)), _p(np.ascontiguousarray(flag, np.uint8)),
                          _p(B), _p(C), M, K, N, threads or THREADS)
    return C


def toCSC(dense):
    dense = _f32(dense)
    rows, cols = dense.shape
    cp = np.empty(cols + 1, dtype=np.int32)
    nnz = lib.orc_toCSC(_p(dense), rows, cols, _p(cp), None, None)
    ri, va = np.empty(nnz, dtype=np.int32), np.empty(nnz, dtype=np.float32)
    lib.orc_toCSC(_p(dense), rows, cols, _p(cp), _p(ri), _p(va))
    return cp, ri, va


def toVectorAwareCSR(dense, threshold=0.0):
    dense = _f32(dense)
    rows, cols = dense.shape
    assert cols % 4 == 0, "toVectorAwareCSR restatement requires cols % 4 == 0"
    n = lib.orc_toVectorAwareCSR(_p(dense), rows, cols, None, None, None, threshold)
    rp, ci, va = np.empty(rows + 1, np.int32), np.empty(n, np.int32), np.empty(n, np.float32)
    lib.orc_toVectorAwareCSR(_p(dense), rows, cols, _p(rp), _p(ci), _p(va), threshold)
    return rp, ci, va


def csr_to_csc(rowptr, colind, val, num_cols):
    rowptr, colind, val = _i32(rowptr), _i32(colind), _f32(val)
    M, nnz = len(rowptr) - 1, len(colind)
    cp, ri = np.empty(num_cols + 1, np.int32), np.empty(nnz, np.int32)
    vt = np.empty(nnz, np.float32) if val is not None else None
    lib.orc_csr_to_csc(_p(rowptr), _p(colind), _p(val), M, num_cols, _p(cp), _p(ri), _p(vt))
    return cp, ri, vt


def densify(rowptr, colind, val, num_cols):
    M = len(rowptr) - 1
    D = np.zeros((M, num_cols), dtype=np.float32)
    rows = np.repeat(np.arange(M), np.diff(rowptr))
    D[rows, colind] = 1.0 if val is None else val
    return D


# ------------------------------------------------------------- FP64
def spmm_f64(rowptr, colind, val, B, Cin=None, threads=None):
    rowptr, colind, val, B = _i32(rowptr), _i32(colind), _f32(val), _f32(B)
    M, N = len(rowptr) - 1, B.shape[1]
    C = np.empty((M, N), dtype=np.float64)
    Cin = _f32(Cin)
    lib.orc_spmm_f64(_p(rowptr), _p(colind), _p(val), _p(B), B.strides[0] // 4, _p(Cin), _p(C), M, N,
                     threads or THREADS)
    return C


def spmv_f64(rowptr, colind, val, x, yin=None, threads=None):
    rowptr, colind, val, x = _i32(rowptr), _i32(colind), _f32(val), _f32(x)
    M = len(rowptr) - 1
    y = np.empty(M, dtype=np.float64)
    yin = _f32(yin)
    lib.orc_spmv_f64(_p(rowptr), _p(colind), _p(val), _p(x), _p(yin), _p(y), M, threads or THREADS)
    return y


def sddmm_f64(rowptr, colind, U, V, scale=None, threads=None):
    rowptr, colind, U, V, scale = _i32(rowptr), _i32(colind), _f32(U), _f32(V), _f32(scale)
    M, kd = len(rowptr) - 1, U.shape[1]
    out = np.empty(len(colind), dtype=np.float64)
    lib.orc_sddmm_f64(_p(rowptr), _p(colind), _p(scale), _p(U), U.strides[0] // 4, _p(V), V.strides[0] // 4, kd,
                      _p(out), M, threads or THREADS)
    return out


def spgemm_f64(a, b, Cin):
    (r1, c1, v1), (r2, c2, v2) = [(_i32(x), _i32(y), _f32(z)) for x, y, z in (a, b)]
    M, N = len(r1) - 1, Cin.shape[1]
    C = np.empty((M, N), dtype=np.float64)
    Cin = _f32(Cin)
    lib.orc_spgemm_f64(_p(r1), _p(c1), _p(v1), _p(r2), _p(c2), _p(v2), _p(Cin), _p(C), M, N)
    return C


# ---------------------------------------------------------- error norms
def row_rel_err(x, ref):
    """Per-row ||x - ref||_2 / max(||ref||_2, tiny) (norm-wise per-row relative error)."""
    x = np.asarray(x, dtype=np.float64).reshape(len(ref), -1)
    ref = np.asarray(ref, dtype=np.float64).reshape(len(ref), -1)
    num = np.linalg.norm(x - ref, axis=1)
    den = np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
    return num / den
