"""ctypes binding of libarmsparse.so (the C ABI declared in include/armsp.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C armsparse_amd/csrc``).  Importing this module without it raises
ImportError: there is no CPU fallback for any operator of this package.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int64, c_uint64, c_void_p

# torch first: it loads the HIP runtime it ships with, and libarmsparse.so's
# libamdhip64.so.7 dependency then resolves to that same runtime.  Loaded the
# other way round the process would hold two HIP runtimes, and launches on
# torch's memory from the second one fail (invalid device).
import torch  # noqa: F401,E402

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libarmsparse.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C armsparse_amd/csrc` "
        "(or __graft_entry__.build()); armsparse_amd has no CPU fallback"
    )

lib = ctypes.CDLL(LIB_PATH)

# status codes (armsp_status)
SUCCESS = 0
ERROR_INVALID_VALUE = 1
ERROR_NOT_SUPPORTED = 2
ERROR_HIP = 3
ERROR_ALLOC = 4
ERROR_NO_DEVICE = 5

# schedules (armsp_schedule)
SCHED_AUTO, SCHED_ROW, SCHED_NNZ, SCHED_MERGE = 0, 1, 2, 3
SCHEDULES = {"auto": SCHED_AUTO, "row": SCHED_ROW, "nnz": SCHED_NNZ, "merge": SCHED_MERGE}

# flags
ACCUMULATE = 0x1
TRUNC4 = 0x2
FUSED_MLA = 0x4
VCSR_BALANCED = 0x8


class Config(ctypes.Structure):
    """armsp_config: schedule and tile shape (zero = automatic)."""

    _fields_ = [
        ("schedule", ctypes.c_int32),
        ("group", ctypes.c_int32),
        ("vec", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("unroll", ctypes.c_int32),
    ]

    @classmethod
    def make(cls, schedule="auto", group=0, vec=0, chunk=0, unroll=0) -> "Config":
        s = SCHEDULES[schedule] if isinstance(schedule, str) else int(schedule)
        return cls(s, int(group), int(vec), int(chunk), int(unroll))

    def as_dict(self) -> dict:
        inv = {v: k for k, v in SCHEDULES.items()}
        return dict(schedule=inv.get(self.schedule, self.schedule), group=self.group, vec=self.vec,
                    chunk=self.chunk, unroll=self.unroll)


class Epilogue(ctypes.Structure):
    """armsp_epilogue: fused column scale, row scale, bias and ReLU of a SpMM."""

    _fields_ = [("col_scale", c_void_p), ("row_scale", c_void_p), ("bias", c_void_p), ("relu", ctypes.c_int32)]


class AdamTensor(ctypes.Structure):
    """armsp_adam_tensor (armsp_optim.h): one float32 tensor of an Adam step (device pointers)."""

    _fields_ = [("param", c_void_p), ("grad", c_void_p), ("exp_avg", c_void_p), ("exp_avg_sq", c_void_p),
                ("step", c_void_p), ("n", ctypes.c_int64), ("weight_decay", ctypes.c_float)]


ADAM_MAX_TENSORS = 16

DTYPE_F32, DTYPE_BF16, DTYPE_FP8_E4M3 = 0, 1, 2

_P = c_void_p  # device or host pointer
_I = c_int
_CFG = POINTER(Config)

_SIGS = {
    "armsp_version": (c_int, []),
    "armsp_status_string": (c_char_p, [_I]),
    "armsp_last_hip_error": (c_int, []),
    "armsp_device_available": (c_int, []),
    "armsp_workspace_reserve": (c_int, [ctypes.c_size_t, POINTER(ctypes.c_size_t)]),
    "armsp_timing_enable": (c_int, [_I]),
    "armsp_timing_reset": (c_int, []),
    "armsp_timing_read": (c_int, [c_char_p, POINTER(c_double), POINTER(c_int64)]),
    "armsp_spmm_f32": (c_int, [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _I, _I, _CFG, _P]),
    "armsp_spmm_f32_noalloc": (c_int, [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _I, _I, _CFG, _P]),
    "armsp_spmm_bf16": (c_int, [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _I, _I, _CFG, _P]),
    "armsp_spmm_fp8": (c_int, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _I, _P, _I, _I, _CFG, _P]),
    "armsp_spmm_ex": (c_int, [_I, _I, _I, _I, _P, _P, _P, _I, _P, _P, _I, _P, _I, _I, POINTER(Epilogue), _CFG,
                              _P]),
    "armsp_spmv_f32": (c_int, [_I, _I, _I, _P, _P, _P, _P, _P, _I, _CFG, _P]),
    "armsp_spmv_vcsr_f32": (c_int, [_I, _I, _P, _P, _P, _P, _P, _I, _P]),
    "armsp_sddmm_csr_f32": (c_int, [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _I, _P, _CFG, _P]),
    "armsp_sddmm_coo_f32": (c_int, [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _I, _P, _P]),
    "armsp_sddmm_csr_bf16": (c_int, [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _I, _P, _CFG, _P]),
    "armsp_spgemm_dense_f32": (c_int, [_I, _I, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _I, _P]),
    "armsp_csr_transpose": (c_int, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "armsp_coo_to_csr": (c_int, [_I, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P]),
    "armsp_csr_to_vcsr_device": (c_int, [_I, _I, _I, _P, _P, _P, c_float, _P, _P, _P, c_int64, _P, _P]),
    "armsp_csr_column_split": (c_int, [_I, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P]),
    "armsp_spmm_max_f32": (c_int, [_I, _I, _I, _I, _P, _P, _P, _P, _I, c_float, _P, _I, _P, _P]),
    "armsp_spmm_datareuse_f32": (c_int, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _I, _P, _I, _I, _P]),
    "armsp_spmm_max_backward_f32": (c_int, [_I, _I, _I, _I, _P, _P, _P, _I, _P, _P, _I, _P]),
    "armsp_gemm_bf16": (c_int, [_I, _I, _I, _P, c_int64, c_int64, _P, c_int64, c_int64, _P, c_int64, _I, _P]),
    "armsp_gemm_bf16_ex": (c_int, [_I, _I, _I, _P, c_int64, c_int64, _P, c_int64, c_int64, _P, _I, c_int64, _I, _P]),
    "armsp_convert_f32_to_bf16": (c_int, [c_int64, _P, _P, _P]),
    "armsp_convert_rows_f32_to_bf16": (c_int, [_I, _I, _P, c_int64, _P, c_int64, _P]),
    "armsp_quantize_fp8_rows": (c_int, [_I, _I, _P, _I, _P, _P, _P]),
    "armsp_dequantize_fp8_rows": (c_int, [_I, _I, _P, _I, _P, _P, _I, _P]),
    "armsp_fill_uniform_f32": (c_int, [c_int64, c_uint64, c_uint64, c_float, c_float, _P, _P]),
    "armsp_fill_normal_f32": (c_int, [c_int64, c_uint64, c_uint64, _P, _P]),
    # learned schedule selector
    "armsp_row_stats": (c_int, [_I, _P, _I, _P, _P, _P]),
    "armsp_sgemm_f32": (c_int, [_I, _I, _I, _P, c_int64, c_int64, _P, c_int64, c_int64, _P, c_int64, c_float, _P,
                                _I, _P]),
    "armsp_softmax_xent": (c_int, [_I, _I, _P, _P, _P, _P, _P, _P]),
    "armsp_relu_backward": (c_int, [c_int64, _P, _P, _P]),
    "armsp_colsum_f32": (c_int, [_I, _I, _P, _P, _P]),
    "armsp_bias_relu_dropout_bf16": (c_int, [_I, _I, _P, c_int64, _P, c_float, _P, _P, _P, _P]),
    "armsp_bias_relu_dropout_backward_bf16": (c_int, [_I, _I, _P, _P, c_float, _P, c_int64, _P, _P]),
    "armsp_adam_step": (c_int, [c_int64, _P, _P, _P, _P, c_float, c_float, c_float, c_float, c_float, _I, _P]),
    "armsp_adam_f32": (c_int, [POINTER(AdamTensor), _I, c_double, c_double, c_double, c_double, _P]),
    "armsp_probe_gather_f32": (c_int, [_I, _P, _P, _I, _I, _I, _I, _I, _P, _P]),
    "armsp_selector_set": (c_int, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "armsp_selector_clear": (c_int, [_I]),
    "armsp_selector_choose": (c_int, [_I, _I, _I, _I, _I, _P, _P, _P]),
    "armsp_mlp_infer": (c_int, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    # host-side generators and format helpers (armsp_gen.h)
    "armsp_rng_u64": (c_uint64, [c_uint64, c_uint64, c_uint64]),
    "armsp_rng_uniform": (c_float, [c_uint64, c_uint64, c_uint64, c_float, c_float]),
    "armsp_host_fill_uniform": (c_int, [c_int64, c_uint64, c_uint64, c_float, c_float, _P]),
    "armsp_host_fill_normal": (c_int, [c_int64, c_uint64, c_uint64, _P]),
    "armsp_gen_uniform_csr": (c_int, [_I, _I, c_int64, c_uint64, _P, _P, _P]),
    "armsp_gen_rmat_csr": (c_int, [_I, c_int64, c_double, c_double, c_double, c_uint64, _P, _P, _P]),
    "armsp_gen_powerlaw_graph": (c_int, [_I, c_int64, c_double, c_uint64, _P, _P]),
    "armsp_gen_gcn_graph": (c_int, [_I, c_int64, c_double, c_uint64, _P, _P]),
    "armsp_gen_corpus_csr": (c_int64, [_I, _I, c_double, c_double, c_uint64, c_int64, _P, _P, _P]),
    "armsp_csr_to_vcsr": (c_int64, [_I, _I, _P, _P, _P, c_float, _P, _P, _P]),
    "armsp_csr_check": (c_int, [_I, _I, c_int64, _P, _P]),
    "armsp_csr_to_datareuse": (c_int, [_I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "armsp_read_mtx": (c_int, [c_char_p, _I, POINTER(c_int), POINTER(c_int), POINTER(c_int64), _P, _P, _P]),
}

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = tuple(_SIGS)


class ArmspError(RuntimeError):
    """A libarmsparse entry point returned a non-success status."""

    def __init__(self, fn: str, status: int):
        msg = lib.armsp_status_string(status).decode()
        hip = lib.armsp_last_hip_error()
        super().__init__(f"{fn}: {msg}" + (f" (hip error {hip})" if status == ERROR_HIP else ""))
        self.status = status


def check(fn: str, status: int) -> None:
    if status != SUCCESS:
        raise ArmspError(fn, status)


def call(name: str, *args) -> None:
    check(name, getattr(lib, name)(*args))


def kernel_time(family: str) -> tuple[float, int]:
    """(total milliseconds, launches) recorded for a kernel family while timing is enabled."""
    ms, n = c_double(0.0), c_int64(0)
    check("armsp_timing_read", lib.armsp_timing_read(family.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value
