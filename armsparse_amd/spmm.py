"""Drop-in for the reference's torch extension module ``spmm``
(gcn/pytorch-custom/spmm.cpp:95-114, 167-172): ``spmm.csr_spmm``.

In the reference op.py the module is JIT-built with
``load(name='spmm', sources=['spmm.cpp'], extra_cflags=['-fopenmp'])``
(gcn/pytorch-custom/op.py:6); replace that line with
``from armsparse_amd import spmm``.
"""
from __future__ import annotations

from typing import Optional

import torch

from . import kernels

__doc__ = "spmm in CSR format on MI355X. csr_spmm is the kernel with edge value."


def csr_spmm(A_rowptr: torch.Tensor, A_colind: torch.Tensor, A_csrVal: Optional[torch.Tensor],
             B: torch.Tensor) -> torch.Tensor:
    """out[M, N] = A·B with A in CSR (M = len(A_rowptr)-1, N = B.shape[1]).

    Same argument contract as the reference: int32 rowptr/colind, float32
    values, contiguous inputs.  The output is zero-initialised (the reference
    allocates it with torch::empty and accumulates into it, spmm.cpp:41,68,
    leaving undefined contents plus A·B).  A_csrVal=None means all-ones edge
    values.  B may also be bfloat16 (fp32 accumulation and output).  The
    result lives on B's device; CPU inputs are staged through the GPU.
    """
    for name, t in (("A_rowptr", A_rowptr), ("A_colind", A_colind), ("B", B)):
        if not t.is_contiguous():
            raise RuntimeError(f"csr_spmm: {name} must be contiguous")
    if A_csrVal is not None and not A_csrVal.is_contiguous():
        raise RuntimeError("csr_spmm: A_csrVal must be contiguous")
    return kernels.spmm(A_rowptr, A_colind, A_csrVal, B)
