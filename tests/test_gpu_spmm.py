"""GPU parity of the SpMM kernels against the oracle.

Bar: the ROW schedule is bit-exact with the restated reference kernels (same
per-element fmaf order); NNZ/MERGE schedules and reduced-precision B are
within 1e-5 norm-wise relative error per row of the restatement and of the
FP64 oracle (for bf16/fp8 B the oracle takes the dequantised B, so only fp32
accumulation error remains)."""
import numpy as np
import pytest
import torch

import oracle
from armsparse_amd import data, kernels
from conftest import rmat_1m_cached, FIX8_ROWSUMS, random_csr

pytestmark = pytest.mark.gpu
TOL = 1e-5


def dev(*arrays):
    return [None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def bits(t):
    return (t.cpu().numpy() if torch.is_tensor(t) else t).view(np.uint32)


def case(msg):
    print(f"[case] {msg}", flush=True)


def check_rows(out, restated, f64, tag=""):
    """Per-row gate: within TOL of FP64, and within TOL of the CPU-order
    restatement beyond that restatement's own distance from FP64 (long rows
    summed sequentially in fp32 drift from FP64 by ~sqrt(n)·eps themselves)."""
    e64 = oracle.row_rel_err(out, f64)
    e_ref = oracle.row_rel_err(out, restated)
    e_ref64 = oracle.row_rel_err(restated, f64)
    print(f"{tag} max err vs fp64 {e64.max():.2e}, vs restated {e_ref.max():.2e}, "
          f"restated vs fp64 {e_ref64.max():.2e}")
    assert e64.max() <= TOL
    assert np.all(e_ref <= TOL + e_ref64)


def test_fix8_all_semantics(fix8):
    rp, ci, va = fix8
    d_rp, d_ci, d_va = dev(rp, ci, va)
    B = torch.ones(7, 6, device="cuda")
    out = kernels.spmm(d_rp, d_ci, d_va, B, config="row")
    assert torch.equal(out[:, 0].cpu(), torch.from_numpy(FIX8_ROWSUMS))
    out = kernels.spmm(d_rp, d_ci, d_va, B, config="merge")
    assert torch.equal(out[:, 0].cpu(), torch.from_numpy(FIX8_ROWSUMS))
    C = torch.zeros(8, 6, device="cuda")
    kernels.spmm(d_rp, d_ci, d_va, B, out=C, accumulate=True, trunc4=True, ncols=4)
    expect = np.zeros((8, 6), np.float32)
    oracle.rowBaseSpmm(rp, ci, va, np.ones((7, 6), np.float32), expect)
    assert np.array_equal(C.cpu().numpy(), expect)


@pytest.mark.parametrize("N", [1, 2, 3, 8, 31, 64, 100, 128, 256, 300])
def test_row_schedule_bitexact(N):
    case(f"row N={N}")
    rp, ci, va = random_csr(600, 500, 0.02, 100 + N, empty_rows=0.1)
    r = np.random.default_rng(N)
    B = r.uniform(-1, 1, (500, N)).astype(np.float32)
    C0 = r.uniform(-1, 1, (600, N)).astype(np.float32)
    d_rp, d_ci, d_va, d_B, d_C = dev(rp, ci, va, B, C0)
    kernels.spmm(d_rp, d_ci, d_va, d_B, out=d_C, accumulate=True, config="row")
    assert np.array_equal(bits(d_C), bits(oracle.rowBaseSpmm3(rp, ci, va, B, C0.copy())))
    out = kernels.spmm(d_rp, d_ci, d_va, d_B, config="row")
    assert np.array_equal(bits(out), bits(oracle.origin(rp, ci, va, B, np.zeros_like(C0))))


@pytest.mark.parametrize("group", [4, 8, 16, 32, 64])
@pytest.mark.parametrize("vec", [1, 2, 4])
def test_row_schedule_tiles_bitexact(group, vec):
    case(f"row G={group} V={vec}")
    rp, ci, va = random_csr(300, 256, 0.05, group * 10 + vec)
    B = np.random.default_rng(1).uniform(-1, 1, (256, 96)).astype(np.float32)
    d_rp, d_ci, d_va, d_B = dev(rp, ci, va, B)
    out = kernels.spmm(d_rp, d_ci, d_va, d_B, config=dict(schedule="row", group=group, vec=vec))
    assert np.array_equal(bits(out), bits(oracle.origin(rp, ci, va, B, np.zeros((300, 96), np.float32))))


def test_reference_tail_semantics():
    """rowBaseSpmm (trunc4 + N%4) and irregular (first 32·floor(N/32) columns) are bit-exact."""
    rp, ci, va = random_csr(200, 150, 0.05, 77)
    r = np.random.default_rng(77)
    for N in (6, 45, 70):
        B = r.uniform(-1, 1, (150, N)).astype(np.float32)
        C0 = r.uniform(-1, 1, (200, N)).astype(np.float32)
        d_rp, d_ci, d_va, d_B, d_C = dev(rp, ci, va, B, C0)
        kernels.spmm(d_rp, d_ci, d_va, d_B, out=d_C, accumulate=True, trunc4=True, ncols=N & ~3)
        assert np.array_equal(bits(d_C), bits(oracle.rowBaseSpmm(rp, ci, va, B, C0.copy())))
        d_C = dev(C0)[0]
        kernels.spmm(d_rp, d_ci, d_va, d_B, out=d_C, accumulate=True, ncols=N & ~31, config="row")
        assert np.array_equal(bits(d_C), bits(oracle.irregular(rp, ci, va, B, C0.copy())))


# MERGE with chunk ≤ 2048 runs the LDS-staged 32-bit-offset kernels — at N=128 fp32
# AUTO (group 0) is the half-wave 256-byte-slice kernel, group=64 the whole-wave one;
# group=1 or a larger chunk runs the 64-bit-address kernel that NNZ always uses
SPLIT_CFGS = ([dict(schedule=s, chunk=c, unroll=u) for s in ("merge", "nnz") for c in (64, 512, 2048)
               for u in (4, 8, 16)]
              + [dict(schedule="merge", chunk=c, unroll=u, group=64) for c in (64, 512, 2048) for u in (4, 8, 16)]
              + [dict(schedule="merge", chunk=c, unroll=u, group=1) for c in (64, 512) for u in (4, 8, 16)]
              + [dict(schedule="merge", chunk=c, unroll=u, group=32) for c in (64, 512, 2048) for u in (4, 8, 16)]
              + [dict(schedule="merge", chunk=c, unroll=u, group=32, vec=4) for c in (64, 512) for u in (4, 8, 16)]
              + [dict(schedule="merge", chunk=4096, unroll=8)])

SLICE = dict(schedule="merge", group=32)  # half a wave per nonzero over 128-byte column slices


@pytest.fixture(scope="module")
def rmat14():
    a = data.rmat_csr(14, 1 << 18, seed=1234)
    return a, dev(a.rowptr, a.colind, a.values)


@pytest.mark.parametrize("cfg", SPLIT_CFGS,
                         ids=lambda c: f"{c['schedule']}-c{c['chunk']}-u{c['unroll']}"
                         + {1: "-wide", 32: "-slice", 64: "-wave"}.get(c.get("group"), "")
                         + (f"-v{c['vec']}" if "vec" in c else ""))
def test_split_schedules_power_law(rmat14, cfg):
    case(f"split {cfg}")
    a, (d_rp, d_ci, d_va) = rmat14
    N = 128
    B = np.random.default_rng(5).uniform(-1, 1, (a.shape[1], N)).astype(np.float32)
    C0 = np.random.default_rng(6).uniform(-1, 1, (a.shape[0], N)).astype(np.float32)
    d_B, d_C = dev(B, C0)
    out = kernels.spmm(d_rp, d_ci, d_va, d_B, config=cfg).cpu().numpy()
    restated = oracle.origin(a.rowptr, a.colind, a.values, B, np.zeros_like(C0))
    f64 = oracle.spmm_f64(a.rowptr, a.colind, a.values, B)
    check_rows(out, restated, f64, str(cfg))
    empty = np.diff(a.rowptr) == 0
    assert not out[empty].any()
    kernels.spmm(d_rp, d_ci, d_va, d_B, out=d_C, accumulate=True, config=cfg)
    f64acc = oracle.spmm_f64(a.rowptr, a.colind, a.values, B, C0)
    assert np.max(oracle.row_rel_err(d_C.cpu().numpy(), f64acc)) <= TOL


@pytest.mark.parametrize("N", [1, 16, 32, 64, 96, 256, 512, 520])
@pytest.mark.parametrize("sched", ["merge", "nnz", "merge-wide", "merge-slice", "merge-slice16B", "merge-wave"])
def test_split_widths(rmat14, N, sched):
    case(f"split width N={N} {sched}")
    a, (d_rp, d_ci, d_va) = rmat14
    B = np.random.default_rng(N).uniform(-1, 1, (a.shape[1], N)).astype(np.float32)
    if sched == "merge-slice16B" and N % 4:
        pytest.skip("16-byte lanes need N % 4 == 0")
    cfg = {"merge-wide": dict(schedule="merge", group=1), "merge-slice": SLICE,
           "merge-slice16B": dict(SLICE, vec=4), "merge-wave": dict(schedule="merge", group=64)}.get(sched, sched)
    out = kernels.spmm(d_rp, d_ci, d_va, dev(B)[0], config=cfg).cpu().numpy()
    ref = oracle.spmm_f64(a.rowptr, a.colind, a.values, B)
    if N >= 16:
        assert np.max(oracle.row_rel_err(out, ref)) <= TOL
    else:  # a handful of columns: gate each element on its dot product's scale Σ|a·b|
        scale = oracle.spmm_f64(a.rowptr, a.colind, np.abs(a.values), np.abs(B))
        assert np.all(np.abs(out - ref) <= TOL * scale + 1e-30)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 7, 8])
@pytest.mark.parametrize("chunk", [256, 512, 1024])
def test_narrow_merge(rmat14, N, chunk):
    """Narrow N with the merge schedule: for N ≤ 4 lanes walk runs of merge items
    (hub rows of the power-law matrix are cut across lanes and waves), from N = 5
    a whole wave per nonzero; overwrite and accumulate."""
    case(f"narrow N={N} chunk={chunk}")
    a, (d_rp, d_ci, d_va) = rmat14
    r = np.random.default_rng(100 + N)
    B = r.uniform(-1, 1, (a.shape[1], N)).astype(np.float32)
    C0 = r.uniform(-1, 1, (a.shape[0], N)).astype(np.float32)
    d_B, d_C = dev(B, C0)
    cfg = dict(schedule="merge", chunk=chunk)
    scale = oracle.spmm_f64(a.rowptr, a.colind, np.abs(a.values), np.abs(B)) + 1e-30
    out = kernels.spmm(d_rp, d_ci, d_va, d_B, config=cfg).cpu().numpy()
    assert np.all(np.abs(out - oracle.spmm_f64(a.rowptr, a.colind, a.values, B)) <= TOL * scale)
    kernels.spmm(d_rp, d_ci, d_va, d_B, out=d_C, accumulate=True, config=cfg)
    ref = oracle.spmm_f64(a.rowptr, a.colind, a.values, B, C0)
    assert np.all(np.abs(d_C.cpu().numpy() - ref) <= TOL * (scale + np.abs(C0)))


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
def test_narrow_merge_low_precision(rmat14, dtype):
    a, (d_rp, d_ci, d_va) = rmat14
    B = torch.from_numpy(np.random.default_rng(21).uniform(-1, 1, (a.shape[1], 4)).astype(np.float32)).cuda()
    if dtype == "bf16":
        Bq, sc = B.to(torch.bfloat16), None
        deq = Bq.float().cpu().numpy()
    else:
        Bq, sc = kernels.quantize_fp8_rows(B)
        deq = (Bq.cpu().view(torch.float8_e4m3fn).float() * sc.cpu()[:, None]).numpy()
    out = kernels.spmm(d_rp, d_ci, d_va, Bq, dense_scale=sc, config="merge").cpu().numpy()
    scale = oracle.spmm_f64(a.rowptr, a.colind, np.abs(a.values), np.abs(deq)) + 1e-30
    assert np.all(np.abs(out - oracle.spmm_f64(a.rowptr, a.colind, a.values, deq)) <= TOL * scale)


def test_value_less_and_empty():
    rp, ci, _ = random_csr(300, 200, 0.05, 9, empty_rows=0.3)
    B = np.random.default_rng(9).uniform(-1, 1, (200, 64)).astype(np.float32)
    d_rp, d_ci, d_B = dev(rp, ci, B)
    ones = np.ones(len(ci), np.float32)
    for cfg in ("row", "merge", "nnz"):
        out = kernels.spmm(d_rp, d_ci, None, d_B, config=cfg).cpu().numpy()
        assert np.max(oracle.row_rel_err(out, oracle.spmm_f64(rp, ci, ones, B))) <= TOL
    # no nonzeros at all: overwrite gives zeros, accumulate leaves C
    z_rp = torch.zeros(301, dtype=torch.int32, device="cuda")
    z_ci = torch.zeros(0, dtype=torch.int32, device="cuda")
    out = kernels.spmm(z_rp, z_ci, None, d_B)
    assert out.shape == (300, 64) and not out.any()
    C = torch.full((300, 64), 3.0, device="cuda")
    kernels.spmm(z_rp, z_ci, None, d_B, out=C, accumulate=True)
    assert bool((C == 3.0).all())


@pytest.mark.parametrize("sched", ["row", "merge", "merge-slice"])
def test_bf16_dense(rmat14, sched):
    a, (d_rp, d_ci, d_va) = rmat14
    B = torch.from_numpy(np.random.default_rng(3).uniform(-1, 1, (a.shape[1], 128)).astype(np.float32))
    Bb = B.to(torch.bfloat16)
    cfg = SLICE if sched == "merge-slice" else sched
    out = kernels.spmm(d_rp, d_ci, d_va, Bb.cuda(), config=cfg).cpu().numpy()
    deq = Bb.float().numpy()
    assert np.max(oracle.row_rel_err(out, oracle.spmm_f64(a.rowptr, a.colind, a.values, deq))) <= TOL
    err32 = np.max(oracle.row_rel_err(out, oracle.spmm_f64(a.rowptr, a.colind, a.values, B.numpy())))
    print(f"bf16 B: max row rel err vs fp32-B FP64 oracle = {err32:.3e}")
    assert err32 < 1e-2


def test_fp8_quantiser_matches_torch():
    B = torch.from_numpy(np.random.default_rng(4).standard_normal((2000, 128)).astype(np.float32))
    B[7] = 0.0  # zero row: scale 1
    B[9, :5] = torch.tensor([1e-8, -3e-6, 448.0, 1e4, -1e4])
    q, sc = kernels.quantize_fp8_rows(B.cuda())
    q, sc = q.cpu(), sc.cpu()
    expect_sc = B.abs().amax(1) / 448.0
    expect_sc[expect_sc == 0] = 1.0
    assert torch.equal(sc, expect_sc)
    expect_q = (B / expect_sc[:, None]).to(torch.float8_e4m3fn).view(torch.uint8)
    assert torch.equal(q, expect_q)


def test_fp8_hardware_decode_matches_torch():
    """All 256 e4m3 codes through the packed fp8 conversion the fp8 SpMM gathers use,
    bit for bit against torch's float8_e4m3fn (NaN codes as NaN); and a scaled matrix."""
    codes = torch.arange(256, dtype=torch.uint8).reshape(4, 64).cuda()
    out = kernels.dequantize_fp8_rows(codes).cpu()
    ref = codes.cpu().view(torch.float8_e4m3fn).float()
    nan = torch.isnan(ref)
    assert torch.equal(torch.isnan(out), nan)
    assert torch.equal(out[~nan].view(torch.int32), ref[~nan].view(torch.int32))
    B = torch.from_numpy(np.random.default_rng(4).standard_normal((300, 77)).astype(np.float32)).cuda()
    q, sc = kernels.quantize_fp8_rows(B)
    deq = kernels.dequantize_fp8_rows(q, sc).cpu()
    assert torch.equal(deq, q.cpu().view(torch.float8_e4m3fn).float() * sc.cpu()[:, None])


@pytest.mark.parametrize("sched", ["row", "merge", "merge-slice"])
def test_fp8_dense(rmat14, sched):
    a, (d_rp, d_ci, d_va) = rmat14
    B = torch.from_numpy(np.random.default_rng(8).uniform(-1, 1, (a.shape[1], 128)).astype(np.float32)).cuda()
    q, sc = kernels.quantize_fp8_rows(B)
    cfg = SLICE if sched == "merge-slice" else sched
    out = kernels.spmm(d_rp, d_ci, d_va, q, dense_scale=sc, config=cfg).cpu().numpy()
    deq = (q.cpu().view(torch.float8_e4m3fn).float() * sc.cpu()[:, None]).numpy()
    assert np.max(oracle.row_rel_err(out, oracle.spmm_f64(a.rowptr, a.colind, a.values, deq))) <= TOL
    err32 = np.max(oracle.row_rel_err(out, oracle.spmm_f64(a.rowptr, a.colind, a.values, B.cpu().numpy())))
    print(f"fp8 B: max row rel err vs fp32-B FP64 oracle = {err32:.3e}")
    assert err32 < 0.1


def test_full_size_rmat_1m_n128():
    """Full size: 1M×1M R-MAT, 2^24 nnz, N=128, every row: MERGE against the
    FP64 oracle and against the CPU-order ROW kernel (bit-exact with the
    restated reference by the tests above)."""
    case("full 1M R-MAT N=128")
    a = rmat_1m_cached()
    d_rp, d_ci, d_va = dev(a.rowptr, a.colind, a.values)
    B = torch.empty((a.shape[1], 128), device="cuda")
    kernels.fill_uniform_(B, data.CFG_SEED + 2, data.STREAM_DENSE)
    merged = kernels.spmm(d_rp, d_ci, d_va, B, config="merge").cpu().numpy()
    rowwise = kernels.spmm(d_rp, d_ci, d_va, B, config="row").cpu().numpy()
    f64 = oracle.spmm_f64(a.rowptr, a.colind, a.values, B.cpu().numpy())
    check_rows(merged, rowwise, f64, "1M R-MAT merge")
    sliced = kernels.spmm(d_rp, d_ci, d_va, B, config=SLICE).cpu().numpy()
    check_rows(sliced, rowwise, f64, "1M R-MAT merge, column slices")
    paired = kernels.spmm(d_rp, d_ci, d_va, B, config=dict(SLICE, vec=4)).cpu().numpy()
    check_rows(paired, rowwise, f64, "1M R-MAT merge, half wave per nonzero, 16-byte lanes")
    wave = kernels.spmm(d_rp, d_ci, d_va, B, config=dict(schedule="merge", group=64)).cpu().numpy()
    check_rows(wave, rowwise, f64, "1M R-MAT merge, whole wave per nonzero")


@pytest.mark.parametrize("sched", ["row", "merge", "nnz", "slice", "narrow"])
@pytest.mark.parametrize("accumulate", [False, True])
def test_epilogue_fusion(rmat14, sched, accumulate):
    """row_scale·(A·diag(col_scale)·B) + bias, ReLU — ReLU on rows cut across waves is applied
    only after their carries are added (hub rows of the power-law matrix span many waves)."""
    case(f"epilogue {sched} acc={accumulate}")
    a, (d_rp, d_ci, d_va) = rmat14
    M, K = a.shape
    r = np.random.default_rng(12)
    W = 3 if sched == "narrow" else 64  # narrow: the N ≤ 8 merge kernel
    B = r.uniform(-1, 1, (K, W)).astype(np.float32)
    cs = r.uniform(0.5, 1.5, K).astype(np.float32)
    rs = r.uniform(0.5, 1.5, M).astype(np.float32)
    bias = r.uniform(-0.5, 0.5, W).astype(np.float32)
    C0 = r.uniform(-1, 1, (M, W)).astype(np.float32)
    d_B, d_cs, d_rs, d_bias, d_C = dev(B, cs, rs, bias, C0)
    cfg = {"row": "row", "slice": dict(SLICE, chunk=256), "narrow": dict(schedule="merge", chunk=256)}.get(
        sched, dict(schedule=sched, chunk=256))
    out = kernels.spmm(d_rp, d_ci, d_va, d_B, out=d_C if accumulate else None, accumulate=accumulate,
                       col_scale=d_cs, row_scale=d_rs, bias=d_bias, relu=True, config=cfg).cpu().numpy()
    prod = oracle.spmm_f64(a.rowptr, a.colind, a.values * cs[a.colind], B)
    ref = rs[:, None].astype(np.float64) * prod + bias + (C0 if accumulate else 0)
    ref = np.maximum(ref, 0)
    scale = rs[:, None] * oracle.spmm_f64(a.rowptr, a.colind, np.abs(a.values * cs[a.colind]), np.abs(B)) + 2
    assert np.all(np.abs(out - ref) <= TOL * scale)
    assert (out >= 0).all()


@pytest.mark.parametrize("sched", [None, "row", "nnz"])
def test_dense_operand_beyond_4gib(sched):
    """B with a row stride of 1100 floats over 2^20 rows (4.4 GB): byte offsets of
    its last rows exceed 32 bits, so every schedule must address it with 64-bit
    offsets by itself (AUTO switches to the 64-bit merge kernel); uniform columns
    reach the far rows.  Against FP64."""
    case(f"B beyond 4 GiB, schedule {sched}")
    K, ld, N = 1 << 20, 1100, 64
    a = data.uniform_csr(1 << 14, K, 1 << 20, seed=41)
    full = torch.empty((K, ld), device="cuda")
    kernels.fill_uniform_(full, 7, 16)
    B = full[:, :N]
    assert B.stride(0) * K * 4 > 1 << 32
    d_rp, d_ci, d_va = dev(a.rowptr, a.colind, a.values)
    out = kernels.spmm(d_rp, d_ci, d_va, B, config=sched).cpu().numpy()
    Bh = B.contiguous().cpu().numpy()
    del full, B
    f64 = oracle.spmm_f64(a.rowptr, a.colind, a.values, Bh)
    bound = oracle.spmm_f64(a.rowptr, a.colind, np.abs(a.values), np.abs(Bh))
    assert np.all(np.abs(out - f64) <= 1e-5 * bound + 1e-30)
