"""GPU parity on the benchmark workloads themselves, at their full sizes:

- the `./SpMM 4096 4096 4096` case: 4096² uniform 1% CSR (167,772 nnz), B 4096×64,
  and its SpMV;
- the 65,536² uniform 0.1% matrix (4,294,967 nnz): SpMV and SDDMM at Kd 32 and 128;
- the 1M×1M R-MAT (2^24 nnz) with B in bf16 and in fp8-e4m3 with per-row scales at
  N = 64, 128 and 256;
- the scaled GCN graph (2.4M nodes, 124M edges + self loops, 100 features, hidden
  256): the aggregation SpMM and training steps, graphed against eager.

Gates: ROW SpMM bit-exact with the restated reference kernel; merge-path results
within 1e-5 norm-wise relative error per row of FP64 (on the dequantised B for
bf16/fp8, with the error against the fp32-B FP64 product printed); element-wise
results (SpMV, SDDMM) within 1e-5 of the scale Σ|a·b| of their dot products.  On
the graph-scale matrices the FP64 products are formed on a seeded sample of rows
that always includes the longest ones."""
import numpy as np
import pytest
import torch

import oracle
from armsparse_amd import data, gcn, kernels
from conftest import rmat_1m_cached

pytestmark = pytest.mark.gpu
TOL = 1e-5


def dev(*arrays):
    return [None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def bits(a):
    return (a.cpu().numpy() if torch.is_tensor(a) else a).view(np.uint32)


def case(msg):
    print(f"[case] {msg}", flush=True)


def sample_rows(rowptr, n_random, n_longest, seed):
    """Sorted row ids: the n_longest longest rows plus n_random seeded random rows."""
    M = len(rowptr) - 1
    lens = np.diff(rowptr)
    longest = np.argsort(lens)[-n_longest:]
    rnd = np.random.default_rng(seed).choice(M, n_random, replace=False)
    return np.unique(np.concatenate([longest, rnd]))


def sub_csr(rowptr, colind, val, rows):
    """The CSR of the selected rows (same columns)."""
    lens = (rowptr[rows + 1] - rowptr[rows]).astype(np.int64)
    rp = np.zeros(len(rows) + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    idx = np.concatenate([np.arange(rowptr[r], rowptr[r + 1]) for r in rows]) if len(rows) else np.zeros(0, int)
    return rp, colind[idx], (val[idx] if val is not None else None)


# ------------------------------------------------------------- 4096² uniform 1%
def test_spmm_4096_n64_row_bitexact_and_auto():
    """The `./SpMM 4096 4096 4096` shape with a dense B (SDMM.cpp's kernel family):
    rowBaseSpmm3 (C += A·B) and origin (C = A·B) bit-exact through the ROW schedule,
    AUTO within 1e-5 per row of FP64."""
    case("4096^2 uniform 1% SpMM N=64")
    A, B, _ = data.spmm_small()
    assert A.nnz == 167_772 and B.shape == (4096, 64)
    C0 = np.random.default_rng(1).uniform(-1, 1, (4096, 64)).astype(np.float32)
    d_rp, d_ci, d_va, d_B, d_C = dev(A.rowptr, A.colind, A.values, B, C0)
    kernels.spmm(d_rp, d_ci, d_va, d_B, out=d_C, accumulate=True, config="row")
    assert np.array_equal(bits(d_C), bits(oracle.rowBaseSpmm3(A.rowptr, A.colind, A.values, B, C0.copy())))
    out = kernels.spmm(d_rp, d_ci, d_va, d_B, config="row")
    assert np.array_equal(bits(out), bits(oracle.origin(A.rowptr, A.colind, A.values, B, np.zeros_like(C0))))
    auto = kernels.spmm(d_rp, d_ci, d_va, d_B).cpu().numpy()
    err = oracle.row_rel_err(auto, oracle.spmm_f64(A.rowptr, A.colind, A.values, B)).max()
    print(f"4096^2 AUTO SpMM: max row rel err vs FP64 {err:.2e}")
    assert err <= TOL


def test_spmv_4096():
    case("4096^2 uniform 1% SpMV")
    A, _, x = data.spmm_small()
    d_rp, d_ci, d_va, d_x = dev(A.rowptr, A.colind, A.values, x)
    y = kernels.spmv(d_rp, d_ci, d_va, d_x).cpu().numpy()
    ref = oracle.spmv_f64(A.rowptr, A.colind, A.values, x)
    bound = oracle.spmv_f64(A.rowptr, A.colind, np.abs(A.values), np.abs(x))
    assert np.all(np.abs(y - ref) <= TOL * bound + 1e-30)


# ------------------------------------------------------- 65,536² uniform 0.1%
@pytest.fixture(scope="module")
def u65k():
    a = data.uniform_65k()
    assert a.nnz == 4_294_967
    return a, dev(a.rowptr, a.colind, a.values)


def test_spmv_65k(u65k):
    case("65536^2 uniform 0.1% SpMV")
    a, (d_rp, d_ci, d_va) = u65k
    x = data.host_uniform(a.shape[1], data.CFG_SEED + 1, data.STREAM_DENSE + 1)
    y = kernels.spmv(d_rp, d_ci, d_va, dev(x)[0]).cpu().numpy()
    ref = oracle.spmv_f64(a.rowptr, a.colind, a.values, x)
    bound = oracle.spmv_f64(a.rowptr, a.colind, np.abs(a.values), np.abs(x))
    assert np.all(np.abs(y - ref) <= TOL * bound + 1e-30)


@pytest.mark.parametrize("kd", [32, 128])
def test_sddmm_65k(u65k, kd):
    """out_ij = A_ij·<U_i, V_j> on the 65K pattern with U, V ~ N(0,1)."""
    case(f"65536^2 uniform 0.1% SDDMM Kd={kd}")
    a, (d_rp, d_ci, d_va) = u65k
    U = torch.empty((a.shape[0], kd), device="cuda")
    V = torch.empty((a.shape[1], kd), device="cuda")
    kernels.fill_normal_(U, data.CFG_SEED + 1, 20)
    kernels.fill_normal_(V, data.CFG_SEED + 1, 21)
    out = kernels.sddmm(d_rp, d_ci, U, V, scale=d_va).cpu().numpy()
    Uh, Vh = U.cpu().numpy(), V.cpu().numpy()
    ref = oracle.sddmm_f64(a.rowptr, a.colind, Uh, Vh, scale=a.values)
    bound = oracle.sddmm_f64(a.rowptr, a.colind, np.abs(Uh), np.abs(Vh), scale=np.abs(a.values))
    err = np.abs(out - ref) / (bound + 1e-30)
    print(f"65K SDDMM Kd={kd}: max |err|/scale {err.max():.2e}")
    assert np.all(np.abs(out - ref) <= TOL * bound + 1e-30)


# ------------------------------------------------- 1M R-MAT, bf16 / fp8 B
@pytest.fixture(scope="module")
def rmat1m():
    a = rmat_1m_cached()
    rows = sample_rows(a.rowptr, 30_000, 2_000, seed=7)
    return a, dev(a.rowptr, a.colind, a.values), rows, sub_csr(a.rowptr, a.colind, a.values, rows)


@pytest.mark.parametrize("N", [64, 128, 256])
@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
def test_rmat_1m_low_precision_b(rmat1m, dtype, N):
    """C = A·B with B stored in bf16 or fp8-e4m3 (+ fp32 row scales), fp32 accumulation,
    AUTO schedule: within 1e-5 per row of FP64 on the dequantised B (32,000 rows incl.
    the 2,000 longest); the error against the fp32-B FP64 product is printed."""
    case(f"1M R-MAT {dtype} B N={N}")
    a, (d_rp, d_ci, d_va), rows, (srp, sci, sva) = rmat1m
    B = torch.empty((a.shape[1], N), device="cuda")
    kernels.fill_uniform_(B, data.CFG_SEED + 2, data.STREAM_DENSE)
    if dtype == "bf16":
        Bq, sc = B.to(torch.bfloat16), None
        deq = Bq.float()
    else:
        Bq, sc = kernels.quantize_fp8_rows(B)
        deq = kernels.dequantize_fp8_rows(Bq, sc)
    out = kernels.spmm(d_rp, d_ci, d_va, Bq, dense_scale=sc)
    got = out[torch.from_numpy(rows).cuda().long()].cpu().numpy()
    del out
    f64 = oracle.spmm_f64(srp, sci, sva, deq.cpu().numpy())
    err = oracle.row_rel_err(got, f64).max()
    err32 = oracle.row_rel_err(got, oracle.spmm_f64(srp, sci, sva, B.cpu().numpy())).max()
    print(f"1M R-MAT {dtype} B N={N}: max row rel err vs FP64 (dequantised B) {err:.2e}, "
          f"vs FP64 (fp32 B) {err32:.2e}")
    assert err <= TOL
    assert err32 < (1e-2 if dtype == "bf16" else 0.1)


# ------------------------------------------------------- scaled GCN graph
@pytest.fixture(scope="module")
def scaled_task():
    torch.manual_seed(0)
    return gcn.make_task("bf16", dropout=0.0, capturable=True, **gcn.SCALED)


def test_scaled_gcn_aggregation(scaled_task):
    """The layer aggregation A·X on the 2.4M-node, 126.4M-entry graph with bf16
    features (100 columns on a 128-column pitch): 4,096 random rows and the 64
    highest-degree ones against FP64."""
    case("scaled GCN aggregation")
    t = scaled_task
    rp, ci, cp, ri, w, wt = t.graph
    out = kernels.spmm(rp, ci, w, t.x)
    rph = rp.cpu().numpy()
    rows = sample_rows(rph, 4096, 64, seed=11)
    srp, sci, sva = sub_csr(rph, ci.cpu().numpy(), w.cpu().numpy(), rows)
    got = out[torch.from_numpy(rows).cuda().long()].cpu().numpy()
    del out
    f64 = oracle.spmm_f64(srp, sci, sva, t.x.float().cpu().numpy())
    err = oracle.row_rel_err(got, f64).max()
    print(f"scaled GCN aggregation: max row rel err vs FP64 {err:.2e}")
    assert err <= TOL


def test_scaled_gcn_training_graphed_equals_eager(scaled_task):
    """Training steps on the scaled graph (bf16, the reference's layer order): losses
    finite and falling, and two replays of the captured step equal the eager steps of
    an identically initialised model (dropout 0)."""
    case("scaled GCN training, graphed vs eager")
    a = scaled_task
    b = gcn.retarget(a, "bf16", capturable=True)
    b.model.load_state_dict(a.model.state_dict())
    step = gcn.GraphedStep(a, warmup=1)  # one eager step of `a`, then the capture
    first = gcn.train_step(b)
    losses = [first]
    for i in range(2):
        step.replay()
        eager = gcn.train_step(b)
        torch.cuda.synchronize()
        assert np.isfinite(eager) and abs(step.loss() - eager) <= 1e-6 * max(1.0, abs(eager)), i
        for (na, pa), (_, pb) in zip(a.model.named_parameters(), b.model.named_parameters()):
            d = (pa - pb).abs().max().item()
            assert d <= 1e-6 * max(1.0, pb.abs().max().item()), (i, na, d)
        losses.append(eager)
    print(f"scaled GCN losses: {losses}")
    assert losses[-1] < losses[0]
