"""Drop-in for the reference's (unbuilt) extension module ``sddmm``
(gcn/pytorch-custom/sddmm.cpp:21-68, kernels in sddmm.cu:7-483):
out[e] = <D1[row(e)], D2[col(e)]> on a COO or CSR pattern."""
from __future__ import annotations

from typing import Optional

import torch

from . import kernels

__doc__ = "SDDMM kernel. Format of COO and CSR are provided."


def coo_sddmm(rowind: torch.Tensor, colind: torch.Tensor, D1: torch.Tensor, D2: torch.Tensor,
              scale: Optional[torch.Tensor] = None) -> torch.Tensor:
    """out[e] = <D1[rowind[e]], D2[colind[e]]> (· scale[e] if given), float32 [nnz]."""
    return kernels.sddmm(rowind, colind, D1, D2, scale=scale, coo=True)


def csr_sddmm(rowptr: torch.Tensor, colind: torch.Tensor, D1: torch.Tensor, D2: torch.Tensor,
              scale: Optional[torch.Tensor] = None) -> torch.Tensor:
    """out[e] = <D1[row(e)], D2[colind[e]]> (· scale[e] if given), float32 [nnz]; D1 has M rows."""
    return kernels.sddmm(rowptr, colind, D1, D2, scale=scale, coo=False)
