"""armsparse_amd — MI355X-native (gfx950) sparse kernels behind the ARMSparse
entry points: CSR SpMM, SDDMM, CSR SpMV, SpGEMM-to-dense and the GCN SpMM op,
plus the GCN training loop's Adam step (``optim.Adam``).

The HIP library ``libarmsparse.so`` (C ABI in include/armsp.h) does all the
work; this package binds it for torch tensors.  Importing fails if the
library has not been built: there is no CPU fallback.
"""
from . import _lib  # noqa: F401  (raises ImportError when libarmsparse.so is missing)
from . import kernels, optim, spmm, sddmm  # noqa: F401
from ._lib import ArmspError, Config  # noqa: F401
from .kernels import (  # noqa: F401
    csr_transpose, gemm_bf16, quantize_fp8_rows, sddmm as sddmm_op, set_selector, spgemm_dense, spmv, to_bf16,
    vcsr_spmv,
)
from .kernels import spmm as spmm_op  # noqa: F401

__version__ = "1.0.0"


def device_available() -> bool:
    """True when a gfx950 device is visible to the library."""
    return bool(_lib.lib.armsp_device_available())
