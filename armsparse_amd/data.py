"""Seeded CSR-native inputs (host generators of libarmsparse, armsp_gen.h)
and the standard workload shapes.

Replaces the reference's time-seeded dense-first generators
(prepareSparse + toCSR, SDMM.cpp:329-370) and the Planetoid loading of the
GCN scripts (gcn/pytorch-custom/gcn_custom.py:19-58), which needs a download.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from ._lib import lib

CFG_SEED = 0xA5A50000  # master seed of workload i is CFG_SEED + i
STREAM_DENSE = 16


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


@dataclass
class CSR:
    """Host CSR matrix (int32 indices, float32 values or None)."""

    rowptr: np.ndarray
    colind: np.ndarray
    values: Optional[np.ndarray]
    shape: tuple

    @property
    def nnz(self) -> int:
        return int(self.colind.size)

    def row_lengths(self) -> np.ndarray:
        return np.diff(self.rowptr)


def host_uniform(n: int, seed: int, stream: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    out = np.empty(int(n), dtype=np.float32)
    _lib.check("armsp_host_fill_uniform", lib.armsp_host_fill_uniform(int(n), seed, stream, lo, hi, _p(out)))
    return out


def host_normal(n: int, seed: int, stream: int) -> np.ndarray:
    out = np.empty(int(n), dtype=np.float32)
    _lib.check("armsp_host_fill_normal", lib.armsp_host_fill_normal(int(n), seed, stream, _p(out)))
    return out


def uniform_csr(M: int, K: int, nnz: int, seed: int, values: bool = True) -> CSR:
    rp = np.empty(M + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    va = np.empty(nnz, np.float32) if values else None
    _lib.check("armsp_gen_uniform_csr", lib.armsp_gen_uniform_csr(M, K, nnz, seed, _p(rp), _p(ci), _p(va)))
    return CSR(rp, ci, va, (M, K))


def rmat_csr(scale: int, nnz: int, seed: int, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             values: bool = True) -> CSR:
    M = 1 << scale
    rp = np.empty(M + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    va = np.empty(nnz, np.float32) if values else None
    _lib.check("armsp_gen_rmat_csr", lib.armsp_gen_rmat_csr(scale, nnz, a, b, c, seed, _p(rp), _p(ci), _p(va)))
    return CSR(rp, ci, va, (M, M))


def powerlaw_graph(n: int, pairs: int, seed: int, gamma: float = 3.0) -> CSR:
    """Symmetric directed adjacency (2·pairs entries, no self loops, no values)."""
    rp = np.empty(n + 1, np.int32)
    ci = np.empty(2 * pairs, np.int32)
    _lib.check("armsp_gen_powerlaw_graph", lib.armsp_gen_powerlaw_graph(n, pairs, gamma, seed, _p(rp), _p(ci)))
    return CSR(rp, ci, None, (n, n))


CORPUS_KINDS = {"uniform": 0, "rmat": 1, "banded": 2, "blockdiag": 3}


def corpus_csr(kind: str, M: int, avg_row: float, param: float, seed: int) -> CSR:
    k = CORPUS_KINDS[kind]
    cap = int(32 * M * max(param, 1)) if k == 3 else int(M * (np.ceil(avg_row) + 1))
    rp = np.empty(M + 1, np.int32)
    ci = np.empty(max(cap, 1), np.int32)
    va = np.empty(max(cap, 1), np.float32)
    n = lib.armsp_gen_corpus_csr(k, M, float(avg_row), float(param), seed, cap, _p(rp), _p(ci), _p(va))
    if n < 0:
        raise RuntimeError(f"armsp_gen_corpus_csr({kind}, M={M}) failed")
    return CSR(rp, ci[:n].copy(), va[:n].copy(), (M, M))


def csr_to_vcsr(a: CSR, threshold: float = 0.0) -> CSR:
    """Zero-padded 1×4-blocked CSR of toVectorAwareCSR (SpMV.cpp:15-38)."""
    M, K = a.shape
    rp = np.empty(M + 1, np.int32)
    n = lib.armsp_csr_to_vcsr(M, K, _p(a.rowptr), _p(a.colind), _p(a.values), threshold, _p(rp), None, None)
    ci, va = np.empty(n, np.int32), np.empty(n, np.float32)
    lib.armsp_csr_to_vcsr(M, K, _p(a.rowptr), _p(a.colind), _p(a.values), threshold, _p(rp), _p(ci), _p(va))
    return CSR(rp, ci, va, (M, (K + 3) // 4 * 4))


def csr_check(a: CSR) -> int:
    M, K = a.shape
    return int(lib.armsp_csr_check(M, K, a.nnz, _p(a.rowptr), _p(a.colind)))


# ------------------------------------------------------------ workloads
def csr_to_datareuse(a: CSR):
    """(rowptr, colind, values, flag uint8): the 2-row DataReuse format of
    toDataReuseFormat (armsp_csr_to_datareuse), same entry count as a."""
    rp = np.empty(a.shape[0] + 1, np.int32)
    ci = np.empty(a.nnz, np.int32)
    va = np.empty(a.nnz, np.float32)
    fl = np.empty(a.nnz, np.uint8)
    st = lib.armsp_csr_to_datareuse(a.shape[0], a.shape[1], _p(a.rowptr), _p(a.colind), _p(a.values), _p(rp),
                                    _p(ci), _p(va), _p(fl))
    if st:
        raise ValueError(f"armsp_csr_to_datareuse: error {st}")
    return rp, ci, va, fl


def read_matrix_market(path: str, mirror: bool = True) -> CSR:
    """Matrix Market coordinate file → canonical CSR (armsp_read_mtx): a later
    entry for a position overwrites an earlier one, zeros are dropped, pattern
    entries are 1.0, symmetric files are mirrored unless mirror=False."""
    flags = 0 if mirror else 0x1  # ARMSP_MTX_STORED_TRIANGLE
    m, k, n = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int64(0)
    p = os.fsencode(path)
    st = lib.armsp_read_mtx(p, flags, ctypes.byref(m), ctypes.byref(k), ctypes.byref(n), None, None, None)
    if st:
        raise ValueError(f"armsp_read_mtx({path!r}): error {st}")
    rowptr = np.empty(m.value + 1, np.int32)
    colind = np.empty(n.value, np.int32)
    values = np.empty(n.value, np.float32)
    st = lib.armsp_read_mtx(p, flags, ctypes.byref(m), ctypes.byref(k), ctypes.byref(n), _p(rowptr), _p(colind),
                            _p(values))
    if st:
        raise ValueError(f"armsp_read_mtx({path!r}): error {st}")
    return CSR(rowptr, colind, values, (m.value, k.value))


def spmm_small(seed: int = CFG_SEED + 0):
    """4096×4096 uniform 1% (167,772 nnz), B 4096×64 U(-1,1), x[4096] U(-1,1)."""
    A = uniform_csr(4096, 4096, 167_772, seed)
    B = host_uniform(4096 * 64, seed, STREAM_DENSE).reshape(4096, 64)
    x = host_uniform(4096, seed, STREAM_DENSE + 1)
    return A, B, x


def rmat_1m(seed: int = CFG_SEED + 1) -> CSR:
    """R-MAT scale 20 (a=.57, b=c=.19, d=.05), 2^24 distinct nonzeros, values U(-1,1)."""
    return rmat_csr(20, 1 << 24, seed)


def uniform_65k(seed: int = CFG_SEED + 1) -> CSR:
    """65,536² uniform 0.1% (4,294,967 nnz)."""
    return uniform_csr(65536, 65536, 4_294_967, seed)


@dataclass
class GCNGraph:
    """Graph arrays in the naming of gcn_custom.py:38-58.

    rowptr/colind/value_csr hold the CSC of the adjacency (from adj.tocsc())
    and colptr/rowind/value_csc its CSR (from adj.tocsr()), exactly as the
    reference script names them; for a symmetric graph both coincide.
    """

    n: int
    rowptr: np.ndarray
    colind: np.ndarray
    value_csr: np.ndarray
    colptr: np.ndarray
    rowind: np.ndarray
    value_csc: np.ndarray


def gcn_graph(n: int = 19_717, pairs: int = 44_324, seed: int = CFG_SEED + 3) -> GCNGraph:
    """Power-law symmetric graph plus self loops, as proc() (gcn_custom.py:38-58) builds from Planetoid."""
    rp = np.empty(n + 1, np.int32)
    c = np.empty(2 * pairs + n, np.int32)
    _lib.check("armsp_gen_gcn_graph", lib.armsp_gen_gcn_graph(n, pairs, 3.0, seed, _p(rp), _p(c)))
    ones = np.ones(c.size, np.float32)
    # the adjacency is symmetric: its CSC equals its CSR
    return GCNGraph(n, rp, c, ones, rp.copy(), c.copy(), ones.copy())
