"""The reference's command-line drivers, rebuilt on libarmsparse (same names,
arguments and output lines), run end to end on the GPU."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "armsparse_amd", "bin")


def run(*args):
    print("[case]", " ".join(args), flush=True)
    r = subprocess.run([os.path.join(BIN, args[0]), *args[1:]], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    return r


@pytest.mark.parametrize("args,keys", [
    (("SDMM", "512", "512", "64", "8", "0.01"), ["toCSRTime:", "Time take:", "blas_timeTake:", "Correct Result"]),
    (("numaAwareSDMM", "256", "256", "32"), ["Running parallelSum", "TimeTake", "sumUpTime", "Correct Result"]),
    (("numaAwareSPDM", "256", "256", "32"), ["TimeTake", "Correct Result"]),
    (("balanceNUMAAwareSPDM", "256", "256", "32"), ["TimeTake", "Correct Result"]),
    (("SpMM", "512", "512", "512", "8"), ["M k N sparity:", "timeTake:", "Correct Result"]),
    (("numaAwareSpMM", "256", "256", "256", "8"), ["compute time", "sumup time", "total", "Correct Result"]),
    (("SpMV", "512", "512", "8"), ["m k ratio:", "time take:", "ToTCSRTime:", "Correct Result"]),
    (("numaAwareSpMV", "512", "512", "8"), ["TimeTake", "Correct Result"]),
])
def test_driver(args, keys):
    """Every driver checks its result against an FP64 product on the host: the
    reference's |Δ| <= 0.1 and a per-row relative error of at most 1e-5."""
    r = run(*args)
    assert r.returncode == 0
    for k in keys:
        assert k in r.stdout
    rel = [ln for ln in r.stdout.splitlines() if ln.startswith("max |C - FP64|")]
    assert len(rel) == 1 and float(rel[0].rsplit(":", 1)[1]) <= 1e-5


def test_numa_driver_phase_times():
    """The NUMA lines carry measured phases: launch (time1), wait (time2) and the
    carry fixup that stands in for sumUp."""
    r = run("numaAwareSDMM", "2048", "2048", "64")
    assert r.returncode == 0
    vals = {}
    for ln in r.stdout.splitlines():
        if ln.startswith("time1 ") or ln.startswith("time2 "):
            vals[ln.split()[0]] = float(ln.split()[1])
        if ln.startswith("--memCopyTime"):
            vals["sumUp"] = float(ln.split("sumUpTime:")[1].split()[0])
    assert vals["time1"] > 0 and vals["time2"] >= 0 and vals["sumUp"] > 0


def test_driver_bad_arguments():
    r = run("SDMM", "512")
    assert r.returncode == 1 and "usage: SDMM" in r.stderr


GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.mark.parametrize("driver,keys", [
    ("suitesparse_sdmm", ["m k n ratio:", "rowBaseSpmmTime", "DataReuseSpmmTime"]),
    ("suitesparse_spmm", ["tocsrTime:", "m k n ratio:", "SpMMTime", "C nnz:"]),
    ("suitesparse_spmv", ["m k:", "time take:"]),
])
@pytest.mark.parametrize("matrix", ["gre_1107.mtx", "1138_bus.mtx"])
def test_suitesparse_driver(driver, keys, matrix):
    """A Matrix Market file from the reference's datasets/ through each
    SuiteSparse driver; the first line is the file's "rows cols nnz"."""
    r = run(driver, "8", os.path.join(GOLDEN, matrix))
    assert r.returncode == 0
    first = r.stdout.splitlines()[0].split()
    assert first == {"gre_1107.mtx": ["1107", "1107", "5664"], "1138_bus.mtx": ["1138", "1138", "2596"]}[matrix]
    for k in keys:
        assert k in r.stdout


def test_suitesparse_driver_missing_file():
    r = run("suitesparse_spmv", "8", os.path.join(GOLDEN, "no_such.mtx"))
    assert r.returncode == 1 and "Matrix Market read error 1" in r.stderr
