"""GPU parity of SpMV (CSR and blocked VCSR), SDDMM, SpGEMM-to-dense, the bf16
MFMA GEMM and the device fills, against the oracle."""
import numpy as np
import pytest
import torch

import oracle
from armsparse_amd import data, kernels
from conftest import rmat_1m_cached, random_csr, read_mtx

pytestmark = pytest.mark.gpu
TOL = 1e-5


def dev(*arrays):
    return [None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def case(msg):
    print(f"[case] {msg}", flush=True)


def abs_bound(rowptr, colind, val, x):
    """Σ|a_i x_i| per row: the scale of the rounding error of a dot product."""
    v = np.ones(len(colind)) if val is None else np.abs(val.astype(np.float64))
    return oracle.spmv_f64(rowptr, colind, v.astype(np.float32), np.abs(x))


# ------------------------------------------------------------------ SpMV
@pytest.fixture(scope="module")
def rmat14():
    a = data.rmat_csr(14, 1 << 18, seed=99)
    return a, dev(a.rowptr, a.colind, a.values)


SPMV_CFGS = [dict(schedule="merge", chunk=c) for c in (256, 512, 1024)] + \
            [dict(schedule="row", group=g) for g in (1, 4, 16, 64)]


@pytest.mark.parametrize("cfg", SPMV_CFGS, ids=lambda c: f"{c['schedule']}-{c.get('chunk', c.get('group'))}")
def test_spmv_csr(rmat14, cfg):
    case(f"spmv {cfg}")
    a, (d_rp, d_ci, d_va) = rmat14
    x = np.random.default_rng(1).uniform(-1, 1, a.shape[1]).astype(np.float32)
    y0 = np.random.default_rng(2).uniform(-1, 1, a.shape[0]).astype(np.float32)
    d_x, d_y = dev(x, y0)
    y = kernels.spmv(d_rp, d_ci, d_va, d_x, config=cfg).cpu().numpy()
    ref = oracle.spmv_f64(a.rowptr, a.colind, a.values, x)
    bound = abs_bound(a.rowptr, a.colind, a.values, x)
    assert np.all(np.abs(y - ref) <= TOL * bound + 1e-30)
    assert not y[np.diff(a.rowptr) == 0].any()
    kernels.spmv(d_rp, d_ci, d_va, d_x, out=d_y, accumulate=True, config=cfg)
    ref = oracle.spmv_f64(a.rowptr, a.colind, a.values, x, y0)
    assert np.all(np.abs(d_y.cpu().numpy() - ref) <= TOL * (bound + np.abs(y0)))


@pytest.mark.parametrize("fused", [False, True])
def test_vcsr_bitexact(fused):
    """sparse_dense_vector drop-in: same four partial sums and (s0+s2)+(s1+s3) order."""
    A, _, x = data.spmm_small()
    v = data.csr_to_vcsr(A)
    y0 = np.random.default_rng(3).uniform(-1, 1, 4096).astype(np.float32)
    d_rp, d_ci, d_va, d_x, d_y = dev(v.rowptr, v.colind, v.values, x, y0)
    kernels.vcsr_spmv(d_rp, d_ci, d_va, d_x, d_y, fused=fused)
    expect = oracle.sparse_dense_vector(x, y0.copy(), v.rowptr, v.colind, v.values, fused=fused)
    assert np.array_equal(d_y.cpu().numpy().view(np.uint32), expect.view(np.uint32))


@pytest.mark.parametrize("fused", [False, True])
def test_vcsr_bitexact_skewed_rows(fused):
    """Rows of thousands of blocks (an R-MAT hub) take the kernel's batched loads and
    its tail: still bit-exact with the restated sparse_dense_vector."""
    case(f"vcsr skewed fused={fused}")
    a = data.rmat_csr(14, 1 << 18, seed=11)
    v = data.csr_to_vcsr(a)
    assert np.diff(v.rowptr).max() > 4 * 512  # rows for the wave-per-row kernel
    x = data.host_uniform(v.shape[1], 5, 1)
    y0 = np.random.default_rng(4).uniform(-1, 1, v.shape[0]).astype(np.float32)
    d_rp, d_ci, d_va, d_x, d_y = dev(v.rowptr, v.colind, v.values, x, y0)
    kernels.vcsr_spmv(d_rp, d_ci, d_va, d_x, d_y, fused=fused)
    expect = oracle.sparse_dense_vector(x, y0.copy(), v.rowptr, v.colind, v.values, fused=fused)
    assert np.array_equal(d_y.cpu().numpy().view(np.uint32), expect.view(np.uint32))


@pytest.mark.parametrize("fused", [False, True])
def test_vcsr_bitexact_more_long_rows_than_waves(fused):
    """More rows of 17 to 1024 blocks than the medium-row walker has waves (one
    block of 4 waves per CU): waves walk several rows each; still bit-exact.  Half
    the rows are empty (short-row walker)."""
    case(f"vcsr many medium rows fused={fused}")
    M, K = 4400, 16384
    rp, ci, va = random_csr(M, K, 0.042, 21, empty_rows=0.5)
    v = data.csr_to_vcsr(data.CSR(rp, ci, va, (M, K)))
    nb = np.diff(v.rowptr) // 4
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    assert ((nb > 16) & (nb <= 1024)).sum() > cus * 4 and (nb == 0).sum() > 0
    x = data.host_uniform(K, 6, 1)
    y0 = np.random.default_rng(5).uniform(-1, 1, M).astype(np.float32)
    d_rp, d_ci, d_va, d_x, d_y = dev(v.rowptr, v.colind, v.values, x, y0)
    kernels.vcsr_spmv(d_rp, d_ci, d_va, d_x, d_y, fused=fused)
    expect = oracle.sparse_dense_vector(x, y0.copy(), v.rowptr, v.colind, v.values, fused=fused)
    assert np.array_equal(d_y.cpu().numpy().view(np.uint32), expect.view(np.uint32))


def _rows_of_lengths(lengths, K, seed):
    """CSR whose row r has lengths[r] distinct uniform columns of [0, K), U(-1,1) values."""
    r = np.random.default_rng(seed)
    cols = [np.sort(r.choice(K, n, replace=False)) for n in lengths]
    rp = np.zeros(len(lengths) + 1, np.int32)
    rp[1:] = np.cumsum(lengths)
    ci = np.concatenate(cols).astype(np.int32) if len(cols) else np.zeros(0, np.int32)
    va = r.uniform(-1, 1, ci.size).astype(np.float32)
    return data.CSR(rp, ci, va, (len(lengths), K))


@pytest.mark.parametrize("fused", [False, True])
def test_vcsr_bitexact_all_three_walkers(fused):
    """Rows for every walker, each class outnumbering its workers: more rows of over
    1024 blocks than long-row workgroups (2 per CU), more of 17-1024 blocks than
    medium-row waves (4 per CU), short and empty rows, shuffled; bit-exact with the
    restated sparse_dense_vector in either reading."""
    case(f"vcsr three walkers fused={fused}")
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    K = 65536
    r = np.random.default_rng(33)
    lengths = np.concatenate([r.integers(5000, 7000, 2 * cus + 40),      # > 1024 blocks of 16384
                              r.integers(70, 1000, 4 * cus + 200),       # 17..1024 blocks
                              r.integers(1, 17, 3000), np.zeros(500, np.int64)])
    r.shuffle(lengths)
    v = data.csr_to_vcsr(_rows_of_lengths(lengths, K, 34))
    nb = np.diff(v.rowptr) // 4
    assert (nb > 1024).sum() > 2 * cus and ((nb > 16) & (nb <= 1024)).sum() > 4 * cus
    assert ((nb > 0) & (nb <= 16)).sum() > 1000 and (nb == 0).sum() > 0
    x = data.host_uniform(K, 7, 1)
    y0 = np.random.default_rng(6).uniform(-1, 1, len(lengths)).astype(np.float32)
    y0[::7] = -0.0  # empty rows must still turn -0 into +0 as the reference's y += 0 does
    d_rp, d_ci, d_va, d_x, d_y = dev(v.rowptr, v.colind, v.values, x, y0)
    kernels.vcsr_spmv(d_rp, d_ci, d_va, d_x, d_y, fused=fused)
    expect = oracle.sparse_dense_vector(x, y0.copy(), v.rowptr, v.colind, v.values, fused=fused)
    assert np.array_equal(d_y.cpu().numpy().view(np.uint32), expect.view(np.uint32))


@pytest.mark.parametrize("fused", [False, True])
def test_vcsr_walker_boundaries(fused):
    """Rows of exactly 0, 1, 15, 16, 17, 1023, 1024, 1025 and more blocks (one nonzero
    per block, at a random offset in it): the limits between the 4-lane, the wave and
    the workgroup walkers; bit-exact."""
    case(f"vcsr boundaries fused={fused}")
    K = 1 << 14
    nbs = [0, 1, 15, 16, 17, 31, 32, 33, 1023, 1024, 1025, 1040, 2048, 4000] * 3
    r = np.random.default_rng(41)
    rows = []
    for nb in nbs:
        blk = np.sort(r.choice(K // 4, nb, replace=False))
        rows.append(4 * blk + r.integers(0, 4, nb))
    rp = np.zeros(len(rows) + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    ci = np.concatenate(rows).astype(np.int32)
    va = r.uniform(-1, 1, ci.size).astype(np.float32)
    v = data.csr_to_vcsr(data.CSR(rp, ci, va, (len(rows), K)))
    assert list(np.diff(v.rowptr) // 4) == nbs
    x = data.host_uniform(K, 8, 1)
    y0 = np.random.default_rng(7).uniform(-1, 1, len(rows)).astype(np.float32)
    d_rp, d_ci, d_va, d_x, d_y = dev(v.rowptr, v.colind, v.values, x, y0)
    kernels.vcsr_spmv(d_rp, d_ci, d_va, d_x, d_y, fused=fused)
    expect = oracle.sparse_dense_vector(x, y0.copy(), v.rowptr, v.colind, v.values, fused=fused)
    assert np.array_equal(d_y.cpu().numpy().view(np.uint32), expect.view(np.uint32))


_VCSR_ENV_CHILD = r"""
import sys
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
import numpy as np, torch
import oracle
from armsparse_amd import data, kernels
a = data.rmat_csr(14, 1 << 18, seed=11)
v = data.csr_to_vcsr(a)
x = data.host_uniform(v.shape[1], 5, 1)
ok = True
for fused in (False, True):
    y0 = np.random.default_rng(4).uniform(-1, 1, v.shape[0]).astype(np.float32)
    t = [torch.from_numpy(np.ascontiguousarray(u)).cuda() for u in (v.rowptr, v.colind, v.values, x, y0)]
    kernels.vcsr_spmv(*t, fused=fused)
    expect = oracle.sparse_dense_vector(x, y0.copy(), v.rowptr, v.colind, v.values, fused=fused)
    ok = ok and bool(np.array_equal(t[4].cpu().numpy().view(np.uint32), expect.view(np.uint32)))
print("bitexact", ok)
"""


@pytest.mark.parametrize("long_blocks", [8, 64])
def test_vcsr_long_threshold_from_the_environment(long_blocks):
    """ARMSP_VCSR_LONG_BLOCKS below and above the 16-block short-row limit: every row
    is still summed once, by one walker, bit-exact (a fresh process, since the value
    is read once per process)."""
    import os
    import subprocess
    import sys

    case(f"vcsr ARMSP_VCSR_LONG_BLOCKS={long_blocks}")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", _VCSR_ENV_CHILD, root], capture_output=True, text=True, timeout=240,
                       env=dict(os.environ, ARMSP_VCSR_LONG_BLOCKS=str(long_blocks)))
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0 and r.stdout.split()[-2:] == ["bitexact", "True"]


def test_vcsr_unaligned_operands_bitexact():
    """Values or x off 16-byte alignment (a sliced view) take the 4-lane kernel for
    every row: same bits as the aligned call."""
    case("vcsr unaligned")
    a = data.rmat_csr(14, 1 << 18, seed=11)
    v = data.csr_to_vcsr(a)
    x = data.host_uniform(v.shape[1], 5, 1)
    y0 = np.random.default_rng(4).uniform(-1, 1, v.shape[0]).astype(np.float32)
    d_rp, d_ci = dev(v.rowptr, v.colind)
    vbuf = torch.zeros(v.nnz + 1, device="cuda")
    vbuf[1:] = torch.from_numpy(v.values).cuda()
    xbuf = torch.zeros(x.size + 1, device="cuda")
    xbuf[1:] = torch.from_numpy(x).cuda()
    d_y = torch.from_numpy(y0.copy()).cuda()
    kernels.vcsr_spmv(d_rp, d_ci, vbuf[1:], xbuf[1:], d_y)
    expect = oracle.sparse_dense_vector(x, y0.copy(), v.rowptr, v.colind, v.values)
    assert np.array_equal(d_y.cpu().numpy().view(np.uint32), expect.view(np.uint32))


def test_vcsr_balanced_schedule():
    """ARMSP_VCSR_BALANCED: the merge path over the padded blocks, y += A·x within
    the SpMV tolerance of FP64 (the reference's order is not kept)."""
    case("vcsr balanced")
    a = data.rmat_csr(14, 1 << 18, seed=11)
    v = data.csr_to_vcsr(a)
    x = data.host_uniform(v.shape[1], 5, 1)
    y0 = np.random.default_rng(4).uniform(-1, 1, v.shape[0]).astype(np.float32)
    d_rp, d_ci, d_va, d_x, d_y = dev(v.rowptr, v.colind, v.values, x, y0)
    kernels.vcsr_spmv(d_rp, d_ci, d_va, d_x, d_y, balanced=True)
    ref = oracle.spmv_f64(v.rowptr, v.colind, v.values, x, y0)
    bound = oracle.spmv_f64(v.rowptr, v.colind, np.abs(v.values), np.abs(x))
    assert np.all(np.abs(d_y.cpu().numpy() - ref) <= TOL * (bound + np.abs(y0)))
    with pytest.raises(RuntimeError):
        kernels.vcsr_spmv(d_rp, d_ci, d_va, d_x, d_y, fused=True, balanced=True)


# ------------------------------------------------- CSR → VCSR on the device
def _vcsr_expected(a, threshold=0.0):
    """toVectorAwareCSR restated on the dense matrix, its columns padded with zeros
    to a multiple of 4 (the blocks past K hold zeros)."""
    k4 = (a.shape[1] + 3) // 4 * 4
    return oracle.toVectorAwareCSR(oracle.densify(a.rowptr, a.colind, a.values, k4), threshold=threshold)


def _with_zeros(a, every=3):
    """a with every `every`-th stored value set to an explicit 0."""
    v = a.values.copy()
    v[::every] = 0.0
    return data.CSR(a.rowptr, a.colind, v, a.shape)


@pytest.mark.parametrize("name", ["uniform", "rmat_skewed", "empty_rows", "k_not_multiple_of_4", "pattern",
                                  "explicit_zeros", "threshold"])
def test_csr_to_vcsr_device_matches_toVectorAwareCSR(name):
    """The device converter gives toVectorAwareCSR's arrays bit for bit: rows of
    every length (R-MAT hub rows), empty rows, K not a multiple of 4, a pattern
    (values 1), explicit zeros (a block of zeros only is not stored) and the
    |v| > 0.1 trigger of numaAwareSpMV."""
    case(f"csr_to_vcsr_device {name}")
    thr = 0.0
    if name == "uniform":
        a = data.uniform_csr(64, 96, 400, seed=4)
    elif name == "rmat_skewed":
        a = data.rmat_csr(12, 1 << 16, seed=6, a=0.7, b=0.1, c=0.1)
    elif name == "empty_rows":
        a = data.CSR(*random_csr(300, 120, 0.05, seed=8, empty_rows=0.3), (300, 120))
        assert (np.diff(a.rowptr) == 0).any()
    elif name == "k_not_multiple_of_4":
        a = data.uniform_csr(50, 37, 300, seed=9)
    elif name == "pattern":
        a = data.uniform_csr(64, 96, 400, seed=10, values=False)
    elif name == "explicit_zeros":
        a = _with_zeros(data.uniform_csr(64, 96, 400, seed=11))
    else:
        a, thr = data.uniform_csr(128, 256, 3000, seed=12), 0.1
    rp, ci, va = _vcsr_expected(a, thr)
    d_rp, d_ci, d_va = dev(a.rowptr, a.colind, a.values)
    g_rp, g_ci, g_va = (t.cpu().numpy() for t in kernels.csr_to_vcsr(d_rp, d_ci, d_va, a.shape[1], threshold=thr))
    assert np.array_equal(g_rp, rp) and np.array_equal(g_ci, ci)
    assert np.array_equal(g_va.view(np.uint32), va.view(np.uint32))
    # the host converter builds the same arrays
    h = data.csr_to_vcsr(a, threshold=thr)
    assert np.array_equal(h.rowptr, g_rp) and np.array_equal(h.colind, g_ci) and np.array_equal(h.values, g_va)


def test_csr_to_vcsr_device_full_rmat_1m():
    """The 1M R-MAT (64.5M padded entries): same arrays as the host converter, and
    the bit-exact SpMV over them gives the same y."""
    case("csr_to_vcsr_device 1M R-MAT")
    a = rmat_1m_cached()
    h = data.csr_to_vcsr(a)
    d_rp, d_ci, d_va = dev(a.rowptr, a.colind, a.values)
    g_rp, g_ci, g_va = kernels.csr_to_vcsr(d_rp, d_ci, d_va, a.shape[1])
    assert np.array_equal(g_rp.cpu().numpy(), h.rowptr)
    assert np.array_equal(g_ci.cpu().numpy(), h.colind)
    assert np.array_equal(g_va.cpu().numpy().view(np.uint32), h.values.view(np.uint32))
    x = data.host_uniform(h.shape[1], 5, 1)
    d_x = dev(x)[0]
    y1 = torch.zeros(a.shape[0], device="cuda")
    y2 = torch.zeros(a.shape[0], device="cuda")
    h_rp, h_ci, h_va = dev(h.rowptr, h.colind, h.values)
    kernels.vcsr_spmv(h_rp, h_ci, h_va, d_x, y1)
    kernels.vcsr_spmv(g_rp, g_ci, g_va, d_x, y2)
    assert torch.equal(y1.view(torch.int32), y2.view(torch.int32))


def test_csr_to_vcsr_device_edges_and_errors():
    """No rows, no entries, and malformed input: columns out of order or out of
    range, rowptr not ending at nnz, and outputs too small are refused."""
    from armsparse_amd import _lib
    import ctypes

    case("csr_to_vcsr_device edges")
    g_rp, g_ci, g_va = kernels.csr_to_vcsr(*dev(np.zeros(1, np.int32), np.zeros(0, np.int32),
                                                np.zeros(0, np.float32)), 8)
    assert g_rp.cpu().tolist() == [0] and g_ci.numel() == 0
    g_rp, g_ci, _ = kernels.csr_to_vcsr(*dev(np.zeros(5, np.int32), np.zeros(0, np.int32), None), 8)
    assert g_rp.cpu().tolist() == [0] * 5 and g_ci.numel() == 0
    bad = [
        (np.array([0, 2, 3]), np.array([2, 0, 1])),  # columns not increasing in row 0
        (np.array([0, 2, 3]), np.array([0, 1, 9])),  # column out of range (K = 8)
        (np.array([0, 2, 2]), np.array([0, 1, 3])),  # rowptr[M] != nnz
        (np.array([0, 3, 2]), np.array([0, 1, 3])),  # decreasing rowptr
    ]
    for rp, ci in bad:
        d = dev(rp.astype(np.int32), ci.astype(np.int32), np.ones(len(ci), np.float32))
        with pytest.raises(RuntimeError):
            kernels.csr_to_vcsr(*d, 8)
    a = data.uniform_csr(16, 32, 40, seed=13)
    d_rp, d_ci, d_va = dev(a.rowptr, a.colind, a.values)
    out_rp = torch.empty(17, dtype=torch.int32, device="cuda")
    out_ci = torch.empty(4, dtype=torch.int32, device="cuda")
    out_va = torch.empty(4, dtype=torch.float32, device="cuda")
    padded = ctypes.c_int64(0)
    st = _lib.lib.armsp_csr_to_vcsr_device(16, 32, a.nnz, d_rp.data_ptr(), d_ci.data_ptr(), d_va.data_ptr(), 0.0,
                                           out_rp.data_ptr(), out_ci.data_ptr(), out_va.data_ptr(), 4,
                                           ctypes.byref(padded), None)
    assert st == _lib.ERROR_INVALID_VALUE and padded.value > 4


# ------------------------------------------------------ column blocks
@pytest.mark.parametrize("nb", [1, 2, 5, 8, 63])
@pytest.mark.parametrize("values", [True, False])
def test_column_split_matches_numpy(nb, values):
    """Block x holds, row by row in the row's order, the entries with columns in
    [x·W, (x+1)·W), W = ceil(K/nb); rowptrs from 0; bit-exact values."""
    case(f"column split nb={nb} values={values}")
    a = data.rmat_csr(12, 1 << 16, seed=7, values=values)
    M, K = a.shape
    W = -(-K // nb)
    rows = np.repeat(np.arange(M), np.diff(a.rowptr))
    assert (np.diff(a.rowptr) == 0).any()
    blocks = kernels.column_blocks(*dev(a.rowptr, a.colind, a.values), K, nb)
    assert len(blocks) == nb
    for x, (rp, ci, va) in enumerate(blocks):
        m = (a.colind >= x * W) & (a.colind < (x + 1) * W)
        exp_rp = np.concatenate([[0], np.cumsum(np.bincount(rows[m], minlength=M))])
        assert np.array_equal(rp.cpu().numpy(), exp_rp) and np.array_equal(ci.cpu().numpy(), a.colind[m])
        if values:
            assert np.array_equal(va.cpu().numpy().view(np.uint32), a.values[m].view(np.uint32))
        else:
            assert va is None


def test_column_split_rejects_non_canonical_rows():
    case("column split errors")
    for rp, ci in ((np.array([0, 2, 3]), np.array([2, 0, 1])), (np.array([0, 2, 3]), np.array([0, 1, 9])),
                   (np.array([0, 2, 2]), np.array([0, 1, 3]))):
        with pytest.raises(RuntimeError):
            kernels.column_blocks(*dev(rp.astype(np.int32), ci.astype(np.int32), np.ones(3, np.float32)), 8, 2)


@pytest.mark.parametrize("n", [1, 3, 4])
def test_spmm_over_column_blocks(n):
    """Σ over column blocks (the first writes, the rest add) with the row scale and
    bias of the fused epilogue: within the FP64 bound of the one-product result."""
    case(f"spmm blocks n={n}")
    a = data.rmat_csr(14, 1 << 18, seed=12)
    M, K = a.shape
    d_rp, d_ci, d_va = dev(a.rowptr, a.colind, a.values)
    B = torch.randn(K, n, device="cuda")
    rs = torch.rand(M, device="cuda") + 0.5
    bias = torch.randn(n, device="cuda")
    for nb in (2, 7):
        blocks = kernels.column_blocks(d_rp, d_ci, d_va, K, nb)
        got = kernels.spmm_blocks(blocks, B, row_scale=rs, bias=bias).cpu().numpy().astype(np.float64)
        Bh = B.cpu().numpy()
        ref = oracle.spmm_f64(a.rowptr, a.colind, a.values, Bh) * rs.cpu().numpy()[:, None].astype(np.float64)
        ref += bias.cpu().numpy()
        bound = oracle.spmm_f64(a.rowptr, a.colind, np.abs(a.values), np.abs(Bh)) * rs.cpu().numpy()[:, None]
        assert np.all(np.abs(got - ref) <= TOL * (bound + np.abs(bias.cpu().numpy())) + 1e-30)


def test_spmv_full_rmat_1m():
    case("spmv full 1M R-MAT")
    a = rmat_1m_cached()
    x = data.host_uniform(a.shape[1], data.CFG_SEED + 1, data.STREAM_DENSE + 1)
    d_rp, d_ci, d_va, d_x = dev(a.rowptr, a.colind, a.values, x)
    y = kernels.spmv(d_rp, d_ci, d_va, d_x).cpu().numpy()
    ref = oracle.spmv_f64(a.rowptr, a.colind, a.values, x)
    assert np.all(np.abs(y - ref) <= TOL * abs_bound(a.rowptr, a.colind, a.values, x) + 1e-30)


# ----------------------------------------------------------------- SDDMM
# schedule "merge" (and AUTO) runs the LDS-staged kernel over merge-path chunks,
# "nnz" the unstaged nnz-split kernel
@pytest.mark.parametrize("kd", [1, 3, 8, 32, 100, 128, 256])
@pytest.mark.parametrize("group", [0, 4, 8, 16, 32])
@pytest.mark.parametrize("sched", ["merge", "nnz", "merge-split"])
def test_sddmm_csr(rmat14, kd, group, sched):
    case(f"sddmm kd={kd} G={group} {sched}")
    a, (d_rp, d_ci, d_va) = rmat14
    r = np.random.default_rng(kd)
    U = r.standard_normal((a.shape[0], kd)).astype(np.float32)
    V = r.standard_normal((a.shape[1], kd)).astype(np.float32)
    d_U, d_V = dev(U, V)
    if sched == "merge-split":  # column halves over the XCDs + combine (odd Kd runs unsplit)
        cfg = dict(schedule="merge", group=group, unroll=2)
    elif group == 0:
        cfg = None if sched == "merge" else dict(schedule="nnz")
    else:
        cfg = dict(schedule=sched, group=group, chunk=512 if sched == "merge" else 0)
    ref = oracle.sddmm_f64(a.rowptr, a.colind, U, V, scale=a.values)
    bound = oracle.sddmm_f64(a.rowptr, a.colind, np.abs(U), np.abs(V), scale=np.abs(a.values))
    out = kernels.sddmm(d_rp, d_ci, d_U, d_V, scale=d_va, config=cfg).cpu().numpy()
    assert np.all(np.abs(out - ref) <= TOL * bound + 1e-30)
    out = kernels.sddmm(d_rp, d_ci, d_U, d_V, config=cfg).cpu().numpy()
    ref1 = oracle.sddmm_f64(a.rowptr, a.colind, U, V)
    assert np.all(np.abs(out - ref1) <= TOL * bound / np.maximum(np.abs(a.values), 1e-30) + 1e-30)


def test_sddmm_coo_and_bf16(rmat14):
    a, (d_rp, d_ci, d_va) = rmat14
    r = np.random.default_rng(11)
    U = r.standard_normal((a.shape[0], 32)).astype(np.float32)
    V = r.standard_normal((a.shape[1], 32)).astype(np.float32)
    rows = np.repeat(np.arange(a.shape[0], dtype=np.int32), np.diff(a.rowptr))
    d_ri, d_U, d_V = dev(rows, U, V)
    ref = oracle.sddmm_f64(a.rowptr, a.colind, U, V)
    bound = oracle.sddmm_f64(a.rowptr, a.colind, np.abs(U), np.abs(V))
    out = kernels.sddmm(d_ri, d_ci, d_U, d_V, coo=True).cpu().numpy()
    assert np.all(np.abs(out - ref) <= TOL * bound + 1e-30)
    Ub, Vb = d_U.to(torch.bfloat16), d_V.to(torch.bfloat16)
    refb = oracle.sddmm_f64(a.rowptr, a.colind, Ub.float().cpu().numpy(), Vb.float().cpu().numpy())
    for cfg in (None, dict(schedule="nnz"), dict(schedule="merge", group=8, vec=1, chunk=2048)):
        out = kernels.sddmm(d_rp, d_ci, Ub, Vb, config=cfg).cpu().numpy()
        assert np.all(np.abs(out - refb) <= TOL * bound * 1.01 + 1e-30)


@pytest.mark.parametrize("kd,split", [(32, False), (128, False), (128, True)])
def test_sddmm_full_rmat_1m(kd, split):
    """The 1M R-MAT at both widths of the workload, and Kd=128 split into column
    halves over the XCDs (then the combine pass)."""
    case(f"sddmm full 1M R-MAT Kd={kd} split={split}")
    a = rmat_1m_cached()
    U = data.host_normal(a.shape[0] * kd, data.CFG_SEED + 1, data.STREAM_DENSE + 2).reshape(-1, kd)
    V = data.host_normal(a.shape[1] * kd, data.CFG_SEED + 1, data.STREAM_DENSE + 3).reshape(-1, kd)
    d_rp, d_ci, d_va, d_U, d_V = dev(a.rowptr, a.colind, a.values, U, V)
    cfg = dict(schedule="merge", unroll=2) if split else None
    out = kernels.sddmm(d_rp, d_ci, d_U, d_V, scale=d_va, config=cfg).cpu().numpy()
    ref = oracle.sddmm_f64(a.rowptr, a.colind, U, V, scale=a.values)
    bound = oracle.sddmm_f64(a.rowptr, a.colind, np.abs(U), np.abs(V), scale=np.abs(a.values))
    assert np.all(np.abs(out - ref) <= TOL * bound + 1e-30)


# ---------------------------------------------------------------- SpGEMM
def test_spgemm_literal_config_bitexact():
    """./SpMM 4096 4096 4096: A and B 4096² at 1%, dense C += A·B, bit-exact with
    the restatement (itself bit-exact with the reference's own source)."""
    case("spgemm 4096^3")
    A = data.uniform_csr(4096, 4096, 167_772, seed=21)
    B = data.uniform_csr(4096, 4096, 167_772, seed=22)
    C0 = data.host_uniform(4096 * 4096, 23, 16).reshape(4096, 4096)
    t = dev(A.rowptr, A.colind, A.values, B.rowptr, B.colind, B.values, C0)
    kernels.spgemm_dense((t[0], t[1], t[2]), (t[3], t[4], t[5]), t[6])
    expect = oracle.sparse_matrix_mul3((A.rowptr, A.colind, A.values), (B.rowptr, B.colind, B.values), C0.copy())
    assert np.array_equal(t[6].cpu().numpy().view(np.uint32), expect.view(np.uint32))


def test_spgemm_wide_tiles():
    a = random_csr(50, 40, 0.2, 31)
    b = random_csr(40, 9000, 0.01, 32)
    C0 = np.zeros((50, 9000), np.float32)
    t = dev(*a, *b, C0)
    kernels.spgemm_dense((t[0], t[1], t[2]), (t[3], t[4], t[5]), t[6])
    expect = oracle.sparse_matrix_mul3(a, b, C0.copy())
    assert np.array_equal(t[6].cpu().numpy(), expect)


def test_spgemm_mtx_fixture():
    rp, ci, va, (M, K) = read_mtx("ibm32.mtx")
    C0 = np.zeros((M, K), np.float32)
    t = dev(rp, ci, va, rp, ci, va, C0)
    kernels.spgemm_dense((t[0], t[1], t[2]), (t[3], t[4], t[5]), t[6])
    expect = oracle.sparse_matrix_mul3((rp, ci, va), (rp, ci, va), C0.copy())
    assert np.array_equal(t[6].cpu().numpy(), expect)


# ------------------------------------------------------------------ GEMM
@pytest.mark.parametrize("m,n,k", [(64, 64, 32), (37, 5, 3), (130, 70, 50), (19717, 64, 500), (500, 64, 19717)])
def test_gemm_bf16_mfma_layout(m, n, k):
    """Checks the MFMA operand/accumulator lane layout against an FP64 GEMM of the bf16 inputs."""
    case(f"gemm {m}x{n}x{k}")
    r = np.random.default_rng(m + n + k)
    a = torch.from_numpy(r.standard_normal((m, k)).astype(np.float32)).to(torch.bfloat16)
    b = torch.from_numpy(r.standard_normal((k, n)).astype(np.float32)).to(torch.bfloat16)
    ref = a.double() @ b.double()
    bound = a.double().abs() @ b.double().abs()
    out = kernels.gemm_bf16(a.cuda(), b.cuda()).cpu().double()
    assert bool(((out - ref).abs() <= TOL * bound + 1e-30).all())
    # transposed operands (strided views) and accumulate
    at = a.t().contiguous().cuda().t()
    bt = b.t().contiguous().cuda().t()
    c = torch.ones(m, n, device="cuda")
    kernels.gemm_bf16(at, bt, out=c, accumulate=True)
    assert bool(((c.cpu().double() - ref - 1).abs() <= TOL * (bound + 1) + 1e-30).all())


@pytest.mark.parametrize("m,n,k,layout", [
    (20000, 256, 100, "nn"), (20011, 250, 99, "nn"), (20000, 3, 256, "nn"), (20000, 64, 37, "nn"),
    (20000, 256, 3, "nt"), (12000, 100, 128, "tn"),              # tall: B resident, A streamed once
    (100, 256, 70000, "tn"), (256, 3, 70001, "tn"), (37, 70, 66000, "nn"),   # reduction over a long K
    (256, 256, 66000, "tn"),                                                  # ... with the largest slabs
    (19717, 64, 500, "nn"), (19717, 64, 500, "nt"), (8000, 32, 1000, "tn"),   # tall, B of up to 1024 k in LDS
    (500, 64, 19717, "tn"), (64, 3, 19717, "tn"), (500, 64, 19717, "nt")])    # split-K tiles: xᵀ·dY at PubMed size
def test_gemm_bf16_tall_and_reduction(m, n, k, layout):
    """The GCN layer's shapes: x·W (tall), the gradient back to x (tall, Wᵀ view) and
    xᵀ·dY (reduction, xᵀ view); against an FP64 GEMM of the same bf16 inputs."""
    case(f"gemm {layout} {m}x{n}x{k}")
    r = np.random.default_rng(m + 3 * n + 7 * k)
    a = torch.from_numpy(r.standard_normal((m, k)).astype(np.float32)).to(torch.bfloat16)
    b = torch.from_numpy(r.standard_normal((k, n)).astype(np.float32)).to(torch.bfloat16)
    ref = a.double() @ b.double()
    bound = a.double().abs() @ b.double().abs()
    da = a.cuda() if layout[0] == "n" else a.t().contiguous().cuda().t()
    db = b.cuda() if layout[1] == "n" else b.t().contiguous().cuda().t()
    out = kernels.gemm_bf16(da, db).cpu().double()
    assert bool(((out - ref).abs() <= TOL * bound + 1e-30).all())
    c = torch.full((m, n), 2.0, device="cuda")
    kernels.gemm_bf16(da, db, out=c, accumulate=True)
    assert bool(((c.cpu().double() - ref - 2).abs() <= TOL * (bound + 2) + 1e-30).all())


@pytest.mark.parametrize("m,n,k", [(19717, 64, 500), (500, 64, 19717), (130, 70, 50)])
def test_gemm_bf16_unaligned_views(m, n, k):
    """Operands at an odd element offset (not 16-byte aligned) and with odd strides
    take the element-load staging; same FP64 bound as the aligned layouts."""
    case(f"gemm unaligned {m}x{n}x{k}")
    r = np.random.default_rng(m + 5 * n + 11 * k)
    a = torch.from_numpy(r.standard_normal((m, k)).astype(np.float32)).to(torch.bfloat16)
    b = torch.from_numpy(r.standard_normal((k, n)).astype(np.float32)).to(torch.bfloat16)
    ref = a.double() @ b.double()
    bound = a.double().abs() @ b.double().abs()
    abuf = torch.zeros(m * (k + 3) + 1, dtype=torch.bfloat16)
    bbuf = torch.zeros((k + 1) * n + 1, dtype=torch.bfloat16)
    abuf[1:].as_strided((m, k), (k + 3, 1)).copy_(a)      # row pitch k+3, offset 1
    bbuf[1:].as_strided((k, n), (1, k + 1)).copy_(b)      # column-major, pitch k+1, offset 1
    da = abuf.cuda()[1:].as_strided((m, k), (k + 3, 1))
    db = bbuf.cuda()[1:].as_strided((k, n), (1, k + 1))
    out = kernels.gemm_bf16(da, db).cpu().double()
    assert bool(((out - ref).abs() <= TOL * bound + 1e-30).all())


@pytest.mark.parametrize("m,n,k", [(20000, 256, 100), (20000, 3, 256), (37, 70, 50), (100, 256, 70000),
                                   (19717, 64, 500), (20011, 250, 99), (4099, 37, 200)])
def test_gemm_bf16_output(m, n, k):
    """bf16 output = the fp32 product rounded to nearest even (tall kernel writes it
    directly, staged as 16-byte rows, other shapes round a workspace); also into rows
    of a pitch that is not a multiple of 8 (the tall kernel's 2-byte stores)."""
    case(f"gemm bf16 out {m}x{n}x{k}")
    r = np.random.default_rng(m + n + k + 1)
    a = torch.from_numpy(r.standard_normal((m, k)).astype(np.float32)).to(torch.bfloat16).cuda()
    b = torch.from_numpy(r.standard_normal((k, n)).astype(np.float32)).to(torch.bfloat16).cuda()
    ref = kernels.gemm_bf16(a, b).to(torch.bfloat16)
    out = kernels.gemm_bf16(a, b, out_dtype=torch.bfloat16)
    assert out.dtype == torch.bfloat16 and torch.equal(out, ref)
    odd = torch.zeros((m, n + 3), dtype=torch.bfloat16, device="cuda")[:, :n]
    kernels.gemm_bf16(a, b, out=odd, out_dtype=torch.bfloat16)
    assert torch.equal(odd, ref)


# ----------------------------------------------------------------- fills
def test_device_uniform_fill_matches_host():
    t = torch.empty(100_003, device="cuda")
    kernels.fill_uniform_(t, 77, 16)
    assert np.array_equal(t.cpu().numpy(), data.host_uniform(100_003, 77, 16))
    kernels.fill_uniform_(t, 77, 16, 0.0, 3.0)
    assert np.array_equal(t.cpu().numpy(), data.host_uniform(100_003, 77, 16, 0.0, 3.0))


def test_bf16_conversion_matches_torch():
    x = torch.from_numpy(np.random.default_rng(0).standard_normal(100_000).astype(np.float32) * 1e3)
    x[:4] = torch.tensor([float("inf"), -0.0, 1e-40, 65504.0])
    y = kernels.to_bf16(x.cuda()).cpu()
    assert torch.equal(y.view(torch.int16), x.to(torch.bfloat16).view(torch.int16))


# ------------------------------------------------------------- transpose
def _check_transpose(rp, ci, va, K):
    d = dev(rp, ci, va)
    cp, ri, vt = kernels.csr_transpose(d[0], d[1], d[2], K)
    ecp, eri, evt = oracle.csr_to_csc(rp, ci, va, K)
    assert np.array_equal(cp.cpu().numpy(), ecp)
    assert np.array_equal(ri.cpu().numpy(), eri)
    assert np.array_equal(vt.cpu().numpy().view(np.uint32), evt.view(np.uint32))
    return cp, ri, vt


@pytest.mark.parametrize("shape,density", [((8, 7), None), ((300, 5), 0.3), ((513, 1000), 0.01),
                                           ((100, 70000), 0.001), ((1, 1), 1.0)])
def test_transpose_stable(fix8, shape, density):
    case(f"transpose {shape}")
    if density is None:
        rp, ci, va = fix8
    else:
        rp, ci, va = random_csr(shape[0], shape[1], density, shape[0] + shape[1])
    _check_transpose(rp, ci, va, shape[1])


def test_transpose_power_law_and_roundtrip(rmat14):
    a, (d_rp, d_ci, d_va) = rmat14
    cp, ri, vt = _check_transpose(a.rowptr, a.colind, a.values, a.shape[1])
    rp2, ci2, v2 = kernels.csr_transpose(cp, ri, vt, a.shape[0])
    assert torch.equal(rp2, d_rp) and torch.equal(ci2, d_ci) and torch.equal(v2, d_va)


def test_transpose_empty_and_valueless():
    rp = torch.zeros(6, dtype=torch.int32, device="cuda")
    ci = torch.zeros(0, dtype=torch.int32, device="cuda")
    cp, ri, vt = kernels.csr_transpose(rp, ci, None, 9)
    assert vt is None and ri.numel() == 0 and not cp.any() and cp.numel() == 10


def test_transpose_full_rmat_1m():
    """Full size (1M R-MAT): exact equality with the stable host transpose, and twice = identity."""
    case("transpose full 1M R-MAT")
    a = rmat_1m_cached()
    cp, ri, vt = _check_transpose(a.rowptr, a.colind, a.values, a.shape[1])
    rp2, ci2, v2 = kernels.csr_transpose(cp, ri, vt, a.shape[0])
    assert np.array_equal(rp2.cpu().numpy(), a.rowptr) and np.array_equal(ci2.cpu().numpy(), a.colind)


# ------------------------------------------------------------- COO → CSR
def _scipy_csr(r, c, v, shape, drop_zeros):
    import scipy.sparse as sp
    a = sp.coo_matrix((v.astype(np.float64), (r, c)), shape=shape).tocsr()  # sums duplicates
    a.sort_indices()
    if drop_zeros:
        a.eliminate_zeros()
    return a.indptr.astype(np.int32), a.indices.astype(np.int32), a.data


@pytest.mark.parametrize("drop_zeros", [False, True])
@pytest.mark.parametrize("m,k,n,seed", [(50, 40, 3000, 1), (1000, 70000, 20000, 2), (1, 1, 5, 3), (70000, 3, 9000, 4)])
def test_coo_to_csr_matches_scipy(m, k, n, seed, drop_zeros):
    """Unordered triplets with many duplicates, cancelling pairs and explicit
    zeros (small-integer values, so every sum is exact): structure and values
    equal scipy's canonical CSR (duplicates summed, indices sorted, zeros
    eliminated on request)."""
    case(f"coo_to_csr {m}x{k} n={n} drop={drop_zeros}")
    r = np.random.default_rng(seed)
    rows = r.integers(0, m, n).astype(np.int32)
    cols = r.integers(0, k, n).astype(np.int32)
    vals = r.integers(-3, 4, n).astype(np.float32)
    rows = np.concatenate([rows, rows[:50]]).astype(np.int32)  # exact cancellations
    cols = np.concatenate([cols, cols[:50]]).astype(np.int32)
    vals = np.concatenate([vals, -vals[:50]]).astype(np.float32)
    erp, eci, eva = _scipy_csr(rows, cols, vals, (m, k), drop_zeros)
    rp, ci, va = kernels.coo_to_csr(*dev(rows, cols, vals), m, k, drop_zeros=drop_zeros)
    assert np.array_equal(rp.cpu().numpy(), erp)
    assert np.array_equal(ci.cpu().numpy(), eci)
    assert np.array_equal(va.cpu().numpy().astype(np.float64), eva)
    # pattern only: the same structure without values (nothing is dropped)
    if not drop_zeros:
        rp2, ci2, va2 = kernels.coo_to_csr(*dev(rows, cols), None, m, k)
        assert va2 is None
        assert np.array_equal(rp2.cpu().numpy(), erp) and np.array_equal(ci2.cpu().numpy(), eci)


def test_coo_to_csr_equals_toCSR_of_the_dense_matrix():
    """Every entry of a dense matrix as a triplet, shuffled, with drop_zeros: the
    result is the reference's toCSR of that matrix (restated in the oracle)."""
    case("coo_to_csr vs toCSR")
    r = np.random.default_rng(9)
    dense = np.where(r.random((37, 53)) < 0.2, r.standard_normal((37, 53)), 0.0).astype(np.float32)
    ii, jj = np.meshgrid(np.arange(37), np.arange(53), indexing="ij")
    perm = r.permutation(dense.size)
    rows, cols, vals = (a.reshape(-1)[perm] for a in (ii.astype(np.int32), jj.astype(np.int32), dense))
    rp, ci, va = kernels.coo_to_csr(*dev(rows, cols, vals), 37, 53, drop_zeros=True)
    erp, eci, eva = oracle.toCSR(dense)
    assert np.array_equal(rp.cpu().numpy(), erp)
    assert np.array_equal(ci.cpu().numpy(), eci)
    assert np.array_equal(va.cpu().numpy().view(np.uint32), eva.view(np.uint32))


def test_coo_to_csr_rebuilds_the_1m_rmat():
    """The 1M R-MAT's triplets in a random order come back as the same CSR, bit
    for bit (no duplicates: the values only move)."""
    case("coo_to_csr 1M R-MAT")
    a = rmat_1m_cached()
    rows = np.repeat(np.arange(a.shape[0], dtype=np.int32), np.diff(a.rowptr))
    perm = np.random.default_rng(5).permutation(a.nnz)
    rp, ci, va = kernels.coo_to_csr(*dev(rows[perm], a.colind[perm], a.values[perm]), a.shape[0], a.shape[1])
    assert np.array_equal(rp.cpu().numpy(), a.rowptr)
    assert np.array_equal(ci.cpu().numpy(), a.colind)
    assert np.array_equal(va.cpu().numpy().view(np.uint32), a.values.view(np.uint32))


def test_coo_to_csr_edges():
    case("coo_to_csr edges")
    e = torch.empty(0, dtype=torch.int32, device="cuda")
    rp, ci, va = kernels.coo_to_csr(e, e, torch.empty(0, device="cuda"), 5, 4)
    assert rp.cpu().tolist() == [0] * 6 and ci.numel() == 0 and va.numel() == 0
    # only the last row has entries; all cancel with drop_zeros
    rows, cols = np.array([4, 4], np.int32), np.array([3, 3], np.int32)
    rp, ci, va = kernels.coo_to_csr(*dev(rows, cols, np.array([1.5, -1.5], np.float32)), 5, 4, drop_zeros=True)
    assert rp.cpu().tolist() == [0] * 6 and ci.numel() == 0
    rp, ci, va = kernels.coo_to_csr(*dev(rows, cols, np.array([1.5, 2.0], np.float32)), 5, 4)
    assert rp.cpu().tolist() == [0, 0, 0, 0, 0, 1] and ci.cpu().tolist() == [3] and va.cpu().tolist() == [3.5]
    # an index out of range is an error, found before anything is written through it
    bad = np.array([0, 5], np.int32)
    with pytest.raises(RuntimeError, match="invalid"):
        kernels.coo_to_csr(*dev(bad, np.array([0, 1], np.int32), np.ones(2, np.float32)), 5, 4)
    with pytest.raises(RuntimeError, match="invalid"):
        kernels.coo_to_csr(*dev(np.array([0, 1], np.int32), np.array([0, -1], np.int32), np.ones(2, np.float32)),
                           5, 4)


# ------------------------------------------------------- max aggregation
def max_fold(rp, ci, va, B, init):
    """Sequential fold of the reference's max reducer (binary_reduce_max.cu:18-24,
    151-166) in float32: acc = acc > x ? acc : x in CSR order; argmax = the
    position whose value was taken last."""
    M, N = len(rp) - 1, B.shape[1]
    out = np.full((M, N), np.float32(init), np.float32)
    arg = np.full((M, N), -1, np.int32)
    for r in range(M):
        for p in range(rp[r], rp[r + 1]):
            x = (np.float32(va[p]) if va is not None else np.float32(1)) * B[ci[p]]
            take = ~(out[r] > x)
            out[r] = np.where(take, x, out[r])
            arg[r] = np.where(take, p, arg[r])
    return out, arg


@pytest.mark.parametrize("m,k,n,density,valued", [(300, 200, 40, 0.05, False), (257, 300, 128, 0.02, True),
                                                  (64, 50, 3, 0.3, True), (100, 100, 130, 0.0, False)])
def test_spmm_max_matches_the_sequential_fold(m, k, n, density, valued):
    case(f"spmm_max {m}x{k} n={n} d={density} valued={valued}")
    rp, ci, va = random_csr(m, k, density, m + n, empty_rows=0.2)
    r = np.random.default_rng(n)
    B = r.integers(-4, 5, (k, n)).astype(np.float32)  # many ties: the later position must win
    va = va if valued else None
    out, arg = kernels.spmm_max(*dev(rp, ci, va, B), return_argmax=True)
    eo, ea = max_fold(rp, ci, va, B, kernels.MAX_INIT)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), eo.view(np.uint32))
    assert np.array_equal(arg.cpu().numpy(), ea)


def test_spmm_max_power_law_and_backward(rmat14):
    """Hub rows of a power-law matrix, init -inf; the backward adds each
    output's gradient (integers: exact in any order) into its winning row."""
    case("spmm_max rmat14 + backward")
    a, _ = rmat14
    rp, ci = a.rowptr, a.colind
    k = len(rp) - 1
    r = np.random.default_rng(3)
    B = r.standard_normal((k, 64)).astype(np.float32)
    out, arg = kernels.spmm_max(*dev(rp, ci, None, B), init=float("-inf"), return_argmax=True)
    eo, ea = max_fold(rp, ci, None, B, -np.inf)
    assert np.array_equal(out.cpu().numpy(), eo)
    assert np.array_equal(arg.cpu().numpy(), ea)
    g = r.integers(-3, 4, (k, 64)).astype(np.float32)
    db = kernels.spmm_max_backward(*dev(ci, None, g), arg, k).cpu().numpy()
    want = np.zeros_like(B)
    rows, cols = np.nonzero(ea >= 0)
    np.add.at(want, (ci[ea[rows, cols]], cols), g[rows, cols])
    assert np.array_equal(db, want)


# --------------------------------------------------------- DataReuse SpMM
@pytest.mark.parametrize("source,n", [("random", 24), ("random", 64), ("odd_rows", 33), ("gre_1107.mtx", 128),
                                      ("1138_bus.mtx", 40), ("rmat14", 64)])
def test_datareuse_spmm_bit_exact(source, n, rmat14):
    """The row-pair kernel against the sequential DataReuseSpmm restatement, bit
    for bit, overwriting and accumulating; and against FP64 A·B."""
    case(f"datareuse {source} n={n}")
    if source == "random":
        rp, ci, va = random_csr(300, 250, 0.06, n)
        shape = (300, 250)
    elif source == "odd_rows":
        rp, ci, va = random_csr(77, 40, 0.5, 3)  # odd M, dense rows: many flagged entries
        shape = (77, 40)
    elif source == "rmat14":
        a, _ = rmat14
        rp, ci, va, shape = a.rowptr, a.colind, a.values, a.shape
    else:
        rp, ci, va, shape = read_mtx(source)
    m, k = shape
    frp, fci, fva, ffl = data.csr_to_datareuse(data.CSR(rp, ci, va, shape))
    r = np.random.default_rng(n)
    B = r.standard_normal((k, n)).astype(np.float32)
    C0 = r.standard_normal((m, n)).astype(np.float32)
    want = oracle.DataReuseSpmm(frp, fci, fva, ffl, B, np.zeros((m, n), np.float32))
    want_acc = oracle.DataReuseSpmm(frp, fci, fva, ffl, B, C0.copy())
    d = dev(frp, fci, fva, ffl, B)
    got = kernels.spmm_datareuse(*d).cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    out = torch.from_numpy(C0.copy()).cuda()
    kernels.spmm_datareuse(*d, out=out, accumulate=True)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want_acc.view(np.uint32))
    ref = oracle.spmm_f64(rp, ci, va, B)
    bound = oracle.spmm_f64(rp, ci, np.abs(va), np.abs(B))
    assert np.all(np.abs(got - ref) <= 1e-5 * bound + 1e-30)


def test_datareuse_shim_host_arrays():
    """The reference-name shim DataReuseSpmm (C++ linkage, host arrays, C += A·B)
    on the golden 8×7 fixture's DataReuse format: C gains the reference's row sums."""
    import ctypes

    from armsparse_amd import _lib
    case("DataReuseSpmm shim")
    from conftest import FIX8_COLIND, FIX8_ROWPTR, FIX8_ROWSUMS, FIX8_VAL
    frp, fci, fva, ffl = oracle.toDataReuseFormat(oracle.densify(FIX8_ROWPTR, FIX8_COLIND, FIX8_VAL, 7))
    fn = _lib.lib["_Z13DataReuseSpmmPiS_PfS0_S0_iiiiPb"]
    fn.restype = None
    B = np.ones((7, 6), np.float32)
    C = np.full((8, 6), 2.0, np.float32)
    fl = ffl.astype(np.bool_)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    fn(p(frp), p(fci), p(fva), p(B), p(C), ctypes.c_int(8), ctypes.c_int(7), ctypes.c_int(6), ctypes.c_int(8), p(fl))
    assert np.array_equal(C, 2.0 + np.repeat(FIX8_ROWSUMS[:, None], 6, axis=1))


def test_sddmm_operands_beyond_4gib():
    """U and V with a row stride of 1100 floats over 2^20 rows (4.4 GB each): the
    staged kernel's 32-bit offsets cannot reach them, so the call takes the
    unstaged kernel with 64-bit addresses.  Against FP64."""
    case("sddmm U, V beyond 4 GiB")
    n, ld, kd = 1 << 20, 1100, 8
    a = data.uniform_csr(n, n, 1 << 20, seed=43)
    Uf = torch.empty((n, ld), device="cuda")
    Vf = torch.empty((n, ld), device="cuda")
    kernels.fill_normal_(Uf, 8, 16)
    kernels.fill_normal_(Vf, 9, 16)
    U, V = Uf[:, :kd], Vf[:, :kd]
    d_rp, d_ci, d_va = dev(a.rowptr, a.colind, a.values)
    out = kernels.sddmm(d_rp, d_ci, U, V, scale=d_va).cpu().numpy()
    Uh, Vh = U.contiguous().cpu().numpy(), V.contiguous().cpu().numpy()
    del Uf, Vf, U, V
    ref = oracle.sddmm_f64(a.rowptr, a.colind, Uh, Vh, scale=a.values)
    bound = oracle.sddmm_f64(a.rowptr, a.colind, np.abs(Uh), np.abs(Vh), scale=np.abs(a.values))
    assert np.all(np.abs(out - ref) <= TOL * bound + 1e-30)


def test_concurrent_calls_from_host_threads(rmat14):
    """Four host threads, each with its own stream, call SpMM (merge path:
    partition and carry workspaces from the shared pool), SpMV and SDDMM
    repeatedly; every result equals the serial one bit for bit (the NUMA
    drivers call the kernels from four pthreads at once)."""
    import threading
    case("concurrent host threads")
    a, (d_rp, d_ci, d_va) = rmat14
    r = np.random.default_rng(77)
    Bs = [torch.from_numpy(r.standard_normal((a.shape[1], 64)).astype(np.float32)).cuda() for _ in range(4)]
    xs = [torch.from_numpy(r.standard_normal(a.shape[1]).astype(np.float32)).cuda() for _ in range(4)]
    U = torch.from_numpy(r.standard_normal((a.shape[0], 32)).astype(np.float32)).cuda()

    def work(i):
        return (kernels.spmm(d_rp, d_ci, d_va, Bs[i]), kernels.spmv(d_rp, d_ci, d_va, xs[i]),
                kernels.sddmm(d_rp, d_ci, U, Bs[i][:, :32].contiguous(), scale=d_va))

    serial = [work(i) for i in range(4)]
    torch.cuda.synchronize()
    results, errors = [[None] * 10 for _ in range(4)], []

    def run(i):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for k in range(10):
                    results[i][k] = work(i)
            s.synchronize()
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for i in range(4):
        for k in range(10):
            for got, want in zip(results[i][k], serial[i]):
                assert torch.equal(got, want), (i, k)
