"""Host-side pieces of bench.py: the committed counter summaries are reported only
while their stamp matches the sources they were measured on, and the helpers that
compute the FP64 error of sampled rows agree with a dense product."""
import json
import os

import numpy as np

import bench


def test_committed_summaries_name_their_sources():
    """profiles/*.json name the kernel and the sources they were measured on, and
    bench.py uses one exactly when its stamp is the hash of those sources now."""
    for name, sources in (("spmm_rmat1m_n128_traffic.json", bench.SPMM_SOURCES),
                          ("gemm_mfma.json", bench.GEMM_SOURCES)):
        rec = json.load(open(os.path.join(bench.ROOT, "profiles", name)))
        assert rec["sources"] == sources and len(rec["source_stamp"]) == 16
        fresh = rec["source_stamp"] == bench.source_stamp(sources)
        assert (bench.stamped_profile(name, sources) is not None) == fresh, name
    t = json.load(open(os.path.join(bench.ROOT, "profiles", "spmm_rmat1m_n128_traffic.json")))
    assert t["kernel"].startswith("k_spmm_slice") and t["hbm_bytes_per_launch"] > t["algorithmic_bytes"]


def test_stale_stamp_is_not_reported(tmp_path, monkeypatch):
    """A summary whose stamp does not match the sources is dropped (figure = None)."""
    prof = tmp_path / "profiles"
    prof.mkdir()
    src = tmp_path / "k.hip"
    src.write_text("kernel v1\n")
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    stamp = bench.source_stamp(["k.hip"])
    (prof / "x.json").write_text(json.dumps({"source_stamp": stamp, "v": 1}))
    assert bench.stamped_profile("x.json", ["k.hip"])["v"] == 1
    src.write_text("kernel v2\n")
    assert bench.stamped_profile("x.json", ["k.hip"]) is None
    assert bench.stamped_profile("missing.json", ["k.hip"]) is None


def test_sampled_row_errors_against_dense_products():
    """spmm_rows_err / spmv_rows_err / sddmm_rows_err are 0 for the exact FP64
    answer and see a planted error in one sampled row."""
    rng = np.random.default_rng(5)
    m, k, n = 40, 30, 6
    dense = np.where(rng.random((m, k)) < 0.2, rng.uniform(-1, 1, (m, k)), 0.0)
    rowptr = np.concatenate([[0], np.cumsum((dense != 0).sum(1))]).astype(np.int32)
    colind = np.nonzero(dense)[1].astype(np.int32)
    val = dense[dense != 0].astype(np.float32)
    B = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    rows = bench.sample_rows(rowptr, 10, 3, seed=2)
    assert np.all(np.diff(rows) > 0) and rows.max() < m
    exact = dense.astype(np.float32).astype(np.float64) @ B.astype(np.float64)
    assert bench.spmm_rows_err(rowptr, colind, val, B, exact[rows], rows) < 1e-12
    bad = exact[rows].copy()
    r = int(np.argmax(np.abs(bad).sum(1)))
    bad[r, 0] += 1e-3 * np.abs(bad[r]).sum()
    assert bench.spmm_rows_err(rowptr, colind, val, B, bad, rows) > 1e-4
    x = B[:, 0].copy()
    y = exact[:, 0]
    assert bench.spmv_rows_err(rowptr, colind, val, x, y[rows], rows) < 1e-12
    U = rng.uniform(-1, 1, (m, 4)).astype(np.float32)
    V = rng.uniform(-1, 1, (k, 4)).astype(np.float32)
    rr = np.repeat(np.arange(m), np.diff(rowptr))
    out = (val.astype(np.float64) * (U[rr].astype(np.float64) * V[colind].astype(np.float64)).sum(1))
    assert bench.sddmm_rows_err(rowptr, colind, val, U, V, out, rows) < 1e-12
