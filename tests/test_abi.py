"""The C-ABI library loads, exports every symbol the public headers declare,
and rejects bad arguments on the host before any GPU work (no GPU needed)."""
import os
import re
import subprocess
import sys
import textwrap

import pytest

from conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "armsparse_amd", "libarmsparse.so")


def _declared_c_functions():
    names = set()
    for h in ("armsp.h", "armsp_gen.h", "armsp_optim.h"):
        src = open(os.path.join(INCLUDE, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(armsp_\w+)\s*\(", src))
    return names


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", LIB], check=True, capture_output=True, text=True)
    return {line.split(" T ", 1)[1].split("(")[0].strip() for line in out.stdout.splitlines() if " T " in line}


def test_every_declared_symbol_is_exported():
    exported = _exported()
    declared = _declared_c_functions()
    assert len(declared) >= 30
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    compat = re.findall(r"^(?:void)\s+(\w+)\(", open(os.path.join(INCLUDE, "armsparse_compat.h")).read(), re.M)
    assert set(compat) == {"rowBaseSpmm", "rowBaseSpmm2", "rowBaseSpmm3", "rowWiseBlocking",
                           "rowWiseRegisterBlocking", "irregular", "origin", "sparse_matrix_mul3",
                           "sparse_dense_vector", "csr_spmm_kernel", "DataReuseSpmm"}
    assert set(compat) <= exported
    # helpers the reference drivers define themselves must not collide
    for helper in ("toCSR", "prepareSparse", "getHighResolutionTime", "num_vector_neon", "toDataReuseFormat"):
        assert helper not in exported


def test_python_binding_covers_header():
    from armsparse_amd import _lib

    assert _declared_c_functions() == set(_lib.EXPORTED)


def test_status_and_invalid_arguments_without_gpu():
    from armsparse_amd import _lib

    lib = _lib.lib
    assert lib.armsp_version() == 100
    assert lib.armsp_status_string(0) == b"success"
    assert lib.armsp_status_string(2) == b"not supported"
    # invalid sizes are rejected before any HIP call
    assert lib.armsp_spmm_f32(-1, 4, 4, 0, None, None, None, None, 4, None, 4, 0, None, None) == 1
    assert lib.armsp_spmm_f32(4, 4, 4, 8, None, None, None, None, 4, None, 4, 0, None, None) == 1
    assert lib.armsp_spmm_f32(0, 4, 4, 0, None, None, None, None, 4, None, 4, 0, None, None) == 0  # empty
    assert lib.armsp_spmv_vcsr_f32(4, 6, None, None, None, None, None, 0, None) == 1  # not a multiple of 4
    assert lib.armsp_sddmm_csr_f32(4, 4, -1, 1, None, None, None, None, 1, None, 1, None, None, None) == 1
    assert lib.armsp_gemm_bf16(4, 4, 4, None, 1, 1, None, 1, 1, None, 2, 0, None) == 1  # ldc < N
    # pitched conversion and the activation backward's output pitch
    assert lib.armsp_convert_rows_f32_to_bf16(4, 4, None, 2, None, 4, None) == 1  # ldx < N
    assert lib.armsp_convert_rows_f32_to_bf16(4, 4, None, 4, None, 3, None) == 1  # ldy < N
    assert lib.armsp_convert_rows_f32_to_bf16(0, 4, None, 4, None, 4, None) == 0  # nothing to do
    assert lib.armsp_convert_rows_f32_to_bf16(4, 4, None, 4, None, 4, None) == 1  # null operands
    assert lib.armsp_bias_relu_dropout_backward_bf16(4, 16, None, None, 0.0, None, 8, None, None) == 1  # ldgy < N
    assert lib.armsp_bias_relu_dropout_backward_bf16(4, 16, None, None, 0.0, None, 20, None, None) == 1  # ldgy % 8
    import ctypes

    # device CSR → VCSR: sizes, a missing count pointer, one of the two outputs, rows without entries
    n = ctypes.c_int64(0)
    p8 = ctypes.c_void_p(8)
    assert lib.armsp_csr_to_vcsr_device(-1, 4, 0, p8, None, None, 0.0, p8, None, None, 0, ctypes.byref(n), None) == 1
    assert lib.armsp_csr_to_vcsr_device(4, 4, 4, p8, p8, None, 0.0, p8, None, None, 0, None, None) == 1
    assert lib.armsp_csr_to_vcsr_device(4, 4, 4, p8, p8, None, 0.0, p8, p8, None, 16, ctypes.byref(n), None) == 1
    assert lib.armsp_csr_to_vcsr_device(0, 4, 3, p8, p8, None, 0.0, p8, None, None, 0, ctypes.byref(n), None) == 1
    # column split: block count 1..63, the host start array, the value output when values are given
    starts = (ctypes.c_int64 * 65)()
    assert lib.armsp_csr_column_split(4, 8, 4, p8, p8, None, 0, p8, p8, None, starts, None) == 1
    assert lib.armsp_csr_column_split(4, 8, 4, p8, p8, None, 64, p8, p8, None, starts, None) == 1
    assert lib.armsp_csr_column_split(4, 8, 4, p8, p8, None, 2, p8, p8, None, None, None) == 1
    assert lib.armsp_csr_column_split(4, 8, 4, p8, p8, p8, 2, p8, p8, None, starts, None) == 1
    cfg = _lib.Config.make("row", group=3)

    # group must be a power of two in [4, 64]: rejected on the host
    assert lib.armsp_spmm_f32(4, 4, 4, 0, ctypes.c_void_p(8), None, None, None, 4, ctypes.c_void_p(8), 4, 0,
                              ctypes.byref(cfg), None) == 1


def test_selector_and_probe_arguments_checked_on_the_host():
    """The selector's install/clear/choose and the gather probe reject bad arguments
    before any device call; an empty probe is a no-op."""
    import ctypes

    import numpy as np

    from armsparse_amd import _lib

    lib = _lib.lib
    F, H, C = 17, 4, 2
    f = lambda n: np.zeros(n, np.float32)  # noqa: E731
    mu, sd, W1, b1, W2, b2, W3, b3 = f(F), np.ones(F, np.float32), f(F * H), f(H), f(H * H), f(H), f(H * C), f(C)
    classes = (_lib.Config * C)(_lib.Config.make("merge"), _lib.Config.make("row", group=8))
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    args = [P(a) for a in (mu, sd, W1, b1, W2, b2, W3, b3)]
    assert lib.armsp_selector_set(2, F, H, C, *args, classes) == 1          # no such op
    assert lib.armsp_selector_set(0, F + 1, H, C, *args, classes) == 1      # 17 features
    assert lib.armsp_selector_set(0, F, 0, C, *args, classes) == 1          # empty hidden layer
    assert lib.armsp_selector_set(0, F, H, C, *args[:-1], None, classes) == 1  # a missing array
    assert lib.armsp_selector_set(0, F, H, C, *args, None) == 1             # no class table
    assert lib.armsp_selector_clear(2) == 1 and lib.armsp_selector_clear(0) == 0
    out = _lib.Config()
    assert lib.armsp_selector_choose(0, 4, 4, 4, 4, None, None, None) == 1    # no output
    assert lib.armsp_selector_choose(3, 4, 4, 4, 4, None, ctypes.byref(out), None) == 1
    # no network installed: the heuristic's all-zero configuration, without a device call
    assert lib.armsp_selector_choose(0, 4, 4, 4, 4, None, ctypes.byref(out), None) == 0
    assert (out.schedule, out.group, out.vec, out.chunk, out.unroll) == (0, 0, 0, 0, 0)
    d = ctypes.c_void_p(16)
    assert lib.armsp_probe_gather_f32(-1, d, d, 8, 128, 128, 32, 2, d, None) == 1
    assert lib.armsp_probe_gather_f32(8, d, d, 8, 64, 128, 32, 2, d, None) == 1    # ldb < N
    assert lib.armsp_probe_gather_f32(8, d, d, 8, 128, 128, 24, 2, d, None) == 1   # group
    assert lib.armsp_probe_gather_f32(8, d, d, 8, 128, 128, 32, 3, d, None) == 1   # vec
    assert lib.armsp_probe_gather_f32(8, d, d, 8, 96, 96, 16, 2, d, None) == 2     # 3 slices: not a divisor of 8
    assert lib.armsp_probe_gather_f32(0, d, d, 8, 128, 128, 32, 2, d, None) == 0   # nothing to gather


def test_adam_arguments_checked_on_the_host():
    """armsp_adam_f32 rejects bad tensor lists and hyper-parameters before any
    device call; an empty list is a no-op."""
    import ctypes

    from armsparse_amd import _lib

    lib = _lib.lib
    d = lambda k: ctypes.c_void_p(256 * (k + 1))  # noqa: E731  (never dereferenced)

    def tensors(*specs):
        arr = (_lib.AdamTensor * max(1, len(specs)))()
        for i, (n, step, wd) in enumerate(specs):
            arr[i] = _lib.AdamTensor(d(4 * i), d(4 * i + 1), d(4 * i + 2), d(4 * i + 3), step, n, wd)
        return arr

    ok = (0.01, 0.9, 0.999, 1e-8)
    assert lib.armsp_adam_f32(tensors(), 0, *ok, None) == 0                                  # nothing to update
    assert lib.armsp_adam_f32(None, 1, *ok, None) == 1
    assert lib.armsp_adam_f32(tensors(), -1, *ok, None) == 1
    many = tensors(*[(8, d(100 + i), 0.0) for i in range(_lib.ADAM_MAX_TENSORS + 1)])
    assert lib.armsp_adam_f32(many, _lib.ADAM_MAX_TENSORS + 1, *ok, None) == 1             # over the limit
    assert lib.armsp_adam_f32(tensors((-1, d(90), 0.0)), 1, *ok, None) == 1                 # negative size
    assert lib.armsp_adam_f32(tensors((8, None, 0.0)), 1, *ok, None) == 1                   # no counter
    assert lib.armsp_adam_f32(tensors((8, d(90), -1e-4)), 1, *ok, None) == 1                # negative decay
    assert lib.armsp_adam_f32(tensors((8, d(90), 0.0), (8, d(90), 0.0)), 2, *ok, None) == 1  # shared counter
    one = tensors((8, d(90), 0.0))
    one[0].grad = None
    assert lib.armsp_adam_f32(one, 1, *ok, None) == 1                                       # no gradient
    for bad in ((-0.01, 0.9, 0.999, 1e-8), (0.01, 1.0, 0.999, 1e-8), (0.01, 0.9, -0.1, 1e-8),
                (0.01, 0.9, 0.999, -1e-8), (float("nan"), 0.9, 0.999, 1e-8)):
        assert lib.armsp_adam_f32(tensors((8, d(90), 0.0)), 1, *bad, None) == 1, bad


def test_no_gpu_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from armsparse_amd import kernels

    rp = torch.tensor([0, 1], dtype=torch.int32)
    ci = torch.tensor([0], dtype=torch.int32)
    with pytest.raises(RuntimeError, match="no ROCm device"):
        kernels.spmm(rp, ci, torch.ones(1), torch.ones(1, 4))


_SHIM_MAIN = textwrap.dedent(r"""
    #include "armsparse_compat.h"
    #include <cstring>
    int main(int argc, char** argv) {
        int rp[] = {0, 1, 0}, ci[] = {0, 0}; float v[] = {1, 1}, B[64] = {0}, C[64] = {0};
        if (!std::strcmp(argv[1], "blocking")) rowWiseBlocking(rp, ci, v, B, C, 7, 2, 8, 1);
        if (!std::strcmp(argv[1], "register")) rowWiseRegisterBlocking(rp, ci, v, B, C, 2, 2, 6, 1);
        if (!std::strcmp(argv[1], "decreasing")) rowBaseSpmm3(rp, ci, v, B, C, 2, 2, 4, 1);
        int rp2[] = {0, 1, 2}, ci2[] = {0, 5};
        if (!std::strcmp(argv[1], "range")) rowBaseSpmm3(rp2, ci2, v, B, C, 2, 2, 4, 1);
        // an odd row of toCSR2 (SDMM.cpp:371-394): columns in decreasing order
        int rp3[] = {0, 1, 3}, ci3[] = {0, 1, 0};
        float v3[] = {1, 2, 3};
        if (!std::strcmp(argv[1], "descending")) rowBaseSpmm3(rp3, ci3, v3, B, C, 2, 2, 4, 1);
        // sparse_matrix_mul3's B with column 1 twice in row 0
        int rpb[] = {0, 2, 2}, cib[] = {1, 1};
        if (!std::strcmp(argv[1], "repeated")) sparse_matrix_mul3(rp3, ci3, v3, rpb, cib, v, C, 2, 2, 2, 1);
        return 0;
    }
""")


def _shim_exe(tmp_path):
    src = tmp_path / "shim.cpp"
    src.write_text(_SHIM_MAIN)
    exe = tmp_path / "shim"
    subprocess.run(["g++", "-O1", f"-I{INCLUDE}", str(src), "-o", str(exe), LIB,
                    f"-Wl,-rpath,{os.path.dirname(LIB)}"], check=True)
    return exe


@pytest.mark.parametrize("case,msg", [("blocking", "requires M % 8 == 0"), ("register", "requires N % 4 == 0"),
                                      ("decreasing", "rowPtr is decreasing"),
                                      ("range", "column index out of range"),
                                      ("repeated", "column index repeated within a row")])
def test_compat_shims_reject_bad_input(tmp_path, case, msg):
    """The reference-name shims validate on the host and exit(EXIT_FAILURE) with a message."""
    r = subprocess.run([str(_shim_exe(tmp_path)), case], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert msg in r.stderr and r.stderr.startswith("armsparse: ")


def test_compat_shims_accept_descending_rows(tmp_path):
    """A toCSR2 row (columns in decreasing order) passes validation, as every reference
    kernel accepts it: without a device the call gets as far as the HIP runtime."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device is present: the GPU test runs this input for real")
    r = subprocess.run([str(_shim_exe(tmp_path)), "descending"], capture_output=True, text=True, timeout=60)
    assert "column" not in r.stderr and "rowPtr" not in r.stderr


_INJECT = r"""
import ctypes, sys
import numpy as np
lib = ctypes.CDLL(sys.argv[1])
class Cfg(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("schedule", "group", "vec", "chunk", "unroll")]
rp = np.array([0, 1, 2, 3, 4], np.int32); ci = np.arange(4, dtype=np.int32)
va = np.ones(4, np.float32); B = np.ones((4, 8), np.float32); C = np.zeros((4, 8), np.float32)
p = lambda a: ctypes.c_void_p(a.ctypes.data)
st = lib.armsp_spmm_f32(4, 4, 8, 4, p(rp), p(ci), p(va), p(B), 8, p(C), 8, 0, ctypes.byref(Cfg(3, 0, 0, 0, 0)), None)
res = ctypes.c_size_t(0)
print(st, lib.armsp_workspace_reserve(1 << 20, ctypes.byref(res)), lib.armsp_last_hip_error())
"""


def test_injected_allocation_failure_without_gpu():
    """ARMSP_FAIL_ALLOC: every device allocation of the library fails before any HIP
    call, so the failure path runs without a GPU — an entry point that needs
    workspace (merge-path SpMM) and the capture-arena reservation return
    ARMSP_ERROR_ALLOC (4) with hipErrorOutOfMemory (2) recorded, and a reference-name
    shim exits with the reference's message format."""
    env = dict(os.environ, ARMSP_FAIL_ALLOC="1")
    r = subprocess.run([sys.executable, "-c", _INJECT, LIB], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["4", "4", "2"]
    # without the flag the same call gets as far as the (absent) device instead
    r = subprocess.run([sys.executable, "-c", _INJECT, LIB], capture_output=True, text=True, timeout=120,
                       env={k: v for k, v in os.environ.items() if k != "ARMSP_FAIL_ALLOC"})
    if not os.path.exists("/dev/kfd"):
        assert r.stdout.split()[0] != "4"


def test_injected_allocation_failure_in_a_shim(tmp_path):
    env = dict(os.environ, ARMSP_FAIL_ALLOC="1")
    r = subprocess.run([str(_shim_exe(tmp_path)), "descending"], capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 1
    assert r.stderr.strip() == "armsparse: rowBaseSpmm3: device allocation failed"


def test_one_hip_runtime_whatever_the_import_order():
    """Importing the package before torch must not map a second HIP runtime next to
    torch's (launches on torch memory from the other one fail)."""
    code = ("import armsparse_amd, torch, re\n"
            "print(len(set(re.findall(r'\\S*libamdhip64\\S*', open('/proc/self/maps').read()))))")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "1"
