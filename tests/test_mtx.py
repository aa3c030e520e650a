"""Matrix Market reader (armsp_read_mtx, host) against scipy.io.mmread on the
reference's own data files (datasets/*.mtx, gcn/data/misc/pubmed.mtx), and the
SuiteSparse drivers' rules on small hand-written files: a later entry for a
position overwrites an earlier one, zeros are dropped (toCSR), pattern entries
are 1.0, symmetric files are mirrored."""
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from armsparse_amd import data
from conftest import GOLDEN


@pytest.mark.parametrize("name", ["cage3.mtx", "cage4.mtx", "ibm32.mtx", "1138_bus.mtx", "gre_1107.mtx",
                                  "pubmed.mtx"])
def test_fixtures_match_scipy(name):
    path = os.path.join(GOLDEN, name)
    a = data.read_matrix_market(path)
    m = sp.csr_matrix(scipy.io.mmread(path))  # mirrors symmetric files
    m.sum_duplicates()  # the fixtures hold no duplicates, so summing = last write
    m.eliminate_zeros()
    m.sort_indices()
    assert a.shape == m.shape
    assert data.csr_check(a) == 0
    assert np.array_equal(a.rowptr, m.indptr)
    assert np.array_equal(a.colind, m.indices)
    assert np.array_equal(a.values, m.data.astype(np.float32))


def test_pubmed_is_the_gcn_graph_shape():
    """The real PubMed citation graph: 19,717 nodes, 44,324 undirected edges
    (88,648 directed entries after mirroring) and 3 self loops, all 1.0."""
    a = data.read_matrix_market(os.path.join(GOLDEN, "pubmed.mtx"))
    assert a.shape == (19717, 19717)
    rows = np.repeat(np.arange(19717), np.diff(a.rowptr))
    assert int((rows == a.colind).sum()) == 3
    assert a.nnz == 88648 + 3
    assert np.all(a.values == 1.0)
    t = sp.csr_matrix((a.values, a.colind, a.rowptr), shape=a.shape)
    assert (t != t.T).nnz == 0


def _write(tmp_path, text):
    p = tmp_path / "m.mtx"
    p.write_text(text)
    return str(p)


def test_driver_rules(tmp_path):
    # last write wins, explicit and overwritten-to-zero entries dropped, blank and comment lines skipped
    p = _write(tmp_path, "%%MatrixMarket matrix coordinate real general\n% c\n3 4 6\n1 1 2.5\n3 4 -1\n\n"
                         "1 1 7\n2 2 0\n3 1 1e-3\n3 4 0\n")
    a = data.read_matrix_market(p)
    assert a.shape == (3, 4)
    assert a.rowptr.tolist() == [0, 1, 1, 2] and a.colind.tolist() == [0, 0]
    assert a.values.tolist() == [7.0, np.float32(1e-3)]
    # without a banner: real general, as the drivers read it
    a = data.read_matrix_market(_write(tmp_path, "2 2 2\n2 1 3\n1 2 4\n"))
    assert a.rowptr.tolist() == [0, 1, 2] and a.colind.tolist() == [1, 0] and a.values.tolist() == [4, 3]


def test_symmetric_pattern_integer_skew(tmp_path):
    sym = "%%MatrixMarket matrix coordinate integer symmetric\n3 3 3\n1 1 5\n3 1 2\n3 2 -4\n"
    a = data.read_matrix_market(_write(tmp_path, sym))
    d = sp.csr_matrix((a.values, a.colind, a.rowptr), shape=a.shape).toarray()
    assert d.tolist() == [[5, 0, 2], [0, 0, -4], [2, -4, 0]]
    a = data.read_matrix_market(_write(tmp_path, sym), mirror=False)  # the stored triangle only
    d = sp.csr_matrix((a.values, a.colind, a.rowptr), shape=a.shape).toarray()
    assert d.tolist() == [[5, 0, 0], [0, 0, 0], [2, -4, 0]]
    a = data.read_matrix_market(_write(tmp_path, "%%MatrixMarket matrix coordinate pattern general\n2 3 2\n1 3\n2 1\n"))
    assert a.values.tolist() == [1.0, 1.0] and a.colind.tolist() == [2, 0]
    a = data.read_matrix_market(_write(tmp_path, "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1.5\n"))
    d = sp.csr_matrix((a.values, a.colind, a.rowptr), shape=a.shape).toarray()
    assert d.tolist() == [[0, -1.5], [1.5, 0]]


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", 2),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", 2),
    ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n2 2 1\n", 3),   # fewer entries than declared
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n", 3),            # value missing
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", 4),          # row out of range
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1\n", 4),          # 0 is not a 1-based index
])
def test_errors(tmp_path, text, code):
    with pytest.raises(ValueError, match=f"error {code}"):
        data.read_matrix_market(_write(tmp_path, text))


def test_missing_file(tmp_path):
    with pytest.raises(ValueError, match="error 1"):
        data.read_matrix_market(str(tmp_path / "nope.mtx"))
