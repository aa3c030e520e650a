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
    (("SpMM", "512", "512", "512", "8"), ["M k N sparity:", "timeTake:"]),
    (("numaAwareSpMM", "256", "256", "256", "8"), ["compute time", "sumup time", "total"]),
    (("SpMV", "512", "512", "8"), ["m k ratio:", "time take:", "ToTCSRTime:"]),
    (("numaAwareSpMV", "512", "512", "8"), ["TimeTake"]),
])
def test_driver(args, keys):
    r = run(*args)
    assert r.returncode == 0
    for k in keys:
        assert k in r.stdout


def test_driver_bad_arguments():
    r = run("SDMM", "512")
    assert r.returncode == 1 and "usage: SDMM" in r.stderr
