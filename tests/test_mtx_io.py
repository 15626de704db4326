"""Matrix Market input (xs_mtx_read, host only -- runs without a GPU) against
scipy.io.mmread and the reference reader's documented behaviour
(hip/Tool/ReadMatrix/csr_mmio_highlevel.h): symmetric expansion, the
diagonal report, and the optional sign flip of general files."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import xsolver_amd
from xsolver_amd import read_mtx


def write(path, text):
    path.write_text(text)
    return path


def as_csr(rp, ci, v):
    n = len(rp) - 1
    return sp.csr_matrix((v, ci, rp), shape=(n, n))


def test_general_unsorted_matches_scipy(tmp_path):
    rng = np.random.default_rng(1)
    A = sp.random(40, 40, density=0.15, random_state=2, format="coo") + sp.eye(40)
    A = A.tocoo()
    perm = rng.permutation(A.nnz)           # entries in random order
    lines = ["%%MatrixMarket matrix coordinate real general", "% comment line",
             f"40 40 {A.nnz}"]
    lines += [f"{A.row[k] + 1} {A.col[k] + 1} {float(A.data[k])!r}" for k in perm]
    p = write(tmp_path / "g.mtx", "\n".join(lines) + "\n")
    rp, ci, v, d = read_mtx(p)
    ref = scipy.io.mmread(str(p)).tocsr()
    ref.sort_indices()
    assert np.array_equal(rp, ref.indptr) and np.array_equal(ci, ref.indices)
    assert np.array_equal(v, ref.data)
    for i in range(40):                       # columns sorted within rows
        assert np.all(np.diff(ci[rp[i]:rp[i + 1]]) > 0)
    assert np.array_equal(d, ref.diagonal())
    # the reference reader negates general files (csr_mmio_highlevel.h:67,178)
    rp2, ci2, v2, d2 = read_mtx(p, negate_general=True)
    assert np.array_equal(v2, -v) and np.array_equal(ci2, ci) and np.array_equal(d2, d)


def test_symmetric_and_pattern_expansion(tmp_path):
    text = """%%MatrixMarket matrix coordinate real symmetric
4 4 6
1 1 4.0
2 1 -1.0
2 2 4.0
3 2 -1.5
3 3 4.0
4 4 2.5
"""
    p = write(tmp_path / "s.mtx", text)
    rp, ci, v, d = read_mtx(p, negate_general=True)   # no flip for symmetric files
    dense = as_csr(rp, ci, v).toarray()
    ref = scipy.io.mmread(str(p)).toarray()
    assert np.array_equal(dense, ref) and np.array_equal(dense, dense.T)
    assert np.array_equal(d, [4.0, 4.0, 4.0, 2.5])
    pat = write(tmp_path / "p.mtx", "%%MatrixMarket matrix coordinate pattern general\n3 3 4\n"
                "1 1\n2 2\n3 3\n1 3\n")
    rp, ci, v, d = read_mtx(pat)
    assert np.array_equal(as_csr(rp, ci, v).toarray(),
                          [[1, 0, 1], [0, 1, 0], [0, 0, 1]])


def test_duplicates_summed_and_errors(tmp_path):
    p = write(tmp_path / "dup.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 3\n"
              "1 1 1.5\n1 1 2.0\n2 2 3.0\n")
    rp, ci, v, _ = read_mtx(p)
    assert list(rp) == [0, 1, 2] and list(v) == [3.5, 3.0]
    bad = write(tmp_path / "rect.mtx", "%%MatrixMarket matrix coordinate real general\n2 3 1\n1 1 1\n")
    with pytest.raises(xsolver_amd.XSolverError) as e:
        read_mtx(bad)
    assert e.value.status == xsolver_amd.XS_ERR_UNSUPPORTED
    oob = write(tmp_path / "oob.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n")
    with pytest.raises(xsolver_amd.XSolverError):
        read_mtx(oob)
    with pytest.raises(xsolver_amd.XSolverError):
        read_mtx(tmp_path / "missing.mtx")
