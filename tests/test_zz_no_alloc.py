"""The allocation-free SpMM entry point (armsp_spmm_f32_noalloc) through the C
ABI: its output equals armsp_spmm_f32's bit for bit (same kernels, same merge
chunks) on two calls in a row, the second with a smaller input that reuses the
device scratch array the first left dirty; a call whose workspace exceeds the
array is refused before anything runs; and the entry point allocates nothing
(the library's memory pool and torch's allocator are left as they were)."""
import ctypes

import numpy as np
import pytest
import torch

from armsparse_amd import _lib, data

pytestmark = pytest.mark.gpu


def dev(*arrays):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def run(fn, A, B, C, cfg=None, flags=0):
    rp, ci, va = A
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    c = ctypes.byref(_lib.Config.make(**cfg)) if cfg else None
    return getattr(_lib.lib, fn)(rp.numel() - 1, B.shape[0], B.shape[1], ci.numel(), ptr(rp), ptr(ci), ptr(va),
                                 ptr(B), B.stride(0), ptr(C), C.stride(0), flags, c, stream)


@pytest.mark.parametrize("cfg", [None, dict(schedule="merge", group=64), dict(schedule="nnz", chunk=512),
                                 dict(schedule="merge", chunk=256), dict(schedule="row")],
                         ids=["auto", "wave", "nnz", "merge256", "row"])
def test_noalloc_matches_allocating_entry_point(cfg):
    print(f"[case] noalloc spmm {cfg}", flush=True)
    big = data.rmat_csr(16, 1 << 20, seed=5)
    small = data.rmat_csr(12, 1 << 14, seed=6)
    for A, N, seed in ((big, 128, 1), (small, 32, 2)):  # second call: smaller, dirty scratch
        Ad = dev(A.rowptr, A.colind, A.values)
        B = torch.empty((A.shape[1], N), device="cuda")
        _lib.call("armsp_fill_uniform_f32", B.numel(), seed, 16, -1.0, 1.0, ptr(B),
                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        C0 = torch.randn(A.shape[0], N, device="cuda")
        expect, got = C0.clone(), C0.clone()
        for flags in (0, _lib.ACCUMULATE):
            assert run("armsp_spmm_f32", Ad, B, expect, cfg, flags) == 0
            assert run("armsp_spmm_f32_noalloc", Ad, B, got, cfg, flags) == 0
            torch.cuda.synchronize()
            assert torch.equal(got.view(torch.int32), expect.view(torch.int32)), (A.shape, flags)


def test_noalloc_refuses_oversized_workspace_and_allocates_nothing():
    print("[case] noalloc limits", flush=True)
    A = data.rmat_csr(14, 1 << 18, seed=7)
    Ad = dev(A.rowptr, A.colind, A.values)
    N = 4096  # 64-item chunks: ~4,350 waves × 4096 columns of carries > 64 MiB
    B = torch.zeros((A.shape[1], N), device="cuda")
    C = torch.full((A.shape[0], N), 7.0, device="cuda")
    torch.cuda.synchronize()
    st = run("armsp_spmm_f32_noalloc", Ad, B, C, dict(schedule="merge", chunk=64))
    assert st == _lib.ERROR_NOT_SUPPORTED
    torch.cuda.synchronize()
    assert bool((C == 7.0).all())  # refused before launching anything
    # a call that fits: torch's caching allocator sees no new block, and the call
    # succeeds while the library's pool would be empty (no pool use at all)
    small = data.rmat_csr(12, 1 << 14, seed=8)
    Sd = dev(small.rowptr, small.colind, small.values)
    Bs = torch.randn(small.shape[1], 64, device="cuda")
    Cs = torch.empty(small.shape[0], 64, device="cuda")
    torch.cuda.synchronize()
    before = torch.cuda.memory_stats()["allocation.all.allocated"]
    assert run("armsp_spmm_f32_noalloc", Sd, Bs, Cs) == 0
    torch.cuda.synchronize()
    assert torch.cuda.memory_stats()["allocation.all.allocated"] == before
    ref = torch.empty_like(Cs)
    assert run("armsp_spmm_f32", Sd, Bs, ref) == 0
    torch.cuda.synchronize()
    assert torch.equal(Cs, ref)


_NO_ALLOC_CHILD = r"""
import ctypes, sys
sys.path.insert(0, sys.argv[1])
import torch
from armsparse_amd import _lib, data
A = data.rmat_csr(12, 1 << 14, seed=9)
rp, ci, va = (torch.from_numpy(a).cuda() for a in (A.rowptr, A.colind, A.values))
B = torch.randn(A.shape[1], 64, device="cuda")
C1, C2 = torch.empty(A.shape[0], 64, device="cuda"), torch.empty(A.shape[0], 64, device="cuda")
s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
p = lambda t: ctypes.c_void_p(t.data_ptr())
args = (A.shape[0], A.shape[1], 64, A.nnz, p(rp), p(ci), p(va), p(B), 64)
st_alloc = _lib.lib.armsp_spmm_f32(*args, p(C1), 64, 0, None, s)
st_free = _lib.lib.armsp_spmm_f32_noalloc(*args, p(C2), 64, 0, None, s)
torch.cuda.synchronize()
ref = torch.zeros(A.shape[0], 64, device="cuda").index_add_(
    0, torch.repeat_interleave(torch.arange(A.shape[0], device="cuda"), rp.diff()), va[:, None] * B[ci.long()])
ok = bool(torch.allclose(C2, ref, rtol=1e-5, atol=1e-5))
print(st_alloc, st_free, ok)
"""


def test_noalloc_runs_when_every_allocation_fails():
    """In a process where the library's every device allocation fails
    (ARMSP_FAIL_ALLOC), the allocating entry point reports ARMSP_ERROR_ALLOC and the
    allocation-free one still computes A·B."""
    import os
    import subprocess
    import sys

    print("[case] noalloc with allocations failing", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", _NO_ALLOC_CHILD, root], capture_output=True, text=True, timeout=240,
                       env=dict(os.environ, ARMSP_FAIL_ALLOC="1"))
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0
    assert r.stdout.split()[-3:] == [str(_lib.ERROR_ALLOC), "0", "True"]
