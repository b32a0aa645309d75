"""MatrixMarket reader (gv_read_matrix_market; reference readFile,
marketReader.cu:3-121): host code, checked on the CPU against scipy's
mmread (a library routine) and against the format's special cases."""
import numpy as np
import pytest

import gv_amd as G


def write(tmp_path, name, text, newline="\n"):
    p = tmp_path / name
    p.write_bytes(text.replace("\n", newline).encode())
    return str(p)


def test_matches_scipy_mmread(tmp_path):
    import scipy.io
    import scipy.sparse
    rng = np.random.default_rng(5)
    n, V = 3000, 700
    r = rng.integers(0, V, n)
    c = rng.integers(0, V, n)
    key = np.unique(r.astype(np.int64) * V + c)
    r, c = (key // V).astype(np.int64), (key % V).astype(np.int64)
    m = scipy.sparse.coo_matrix((rng.random(len(r)), (r, c)), shape=(V, V))
    p = str(tmp_path / "g.mtx")
    scipy.io.mmwrite(p, m)
    Vr, s, d, sym = G.read_matrix_market(p)
    ref = scipy.io.mmread(p).tocoo()
    assert Vr == V and not sym
    got = np.sort(s.astype(np.int64) * V + d)
    exp = np.sort(ref.row.astype(np.int64) * V + ref.col)
    assert np.array_equal(got, exp)


def test_pattern_symmetric_comments_blank_lines_crlf_tabs(tmp_path):
    text = ("%%MatrixMarket matrix coordinate pattern symmetric\n"
            "% a comment\n"
            "\n"
            "5 5 4\n"
            "2 1\n"
            "% comment between entries\n"
            "3\t1\n"
            "  5 4  \n"
            "\n"
            "4 2")  # last line without a newline
    for nl in ("\n", "\r\n"):
        V, s, d, sym = G.read_matrix_market(write(tmp_path, "p.mtx", text, nl))
        assert V == 5 and sym
        assert s.tolist() == [1, 2, 4, 3] and d.tolist() == [0, 0, 3, 1]


def test_values_are_ignored_and_rectangular_uses_max_dim(tmp_path):
    text = ("%%MatrixMarket matrix coordinate real general\n"
            "3 6 3\n"
            "1 6 0.5\n"
            "3 1 -2e3\n"
            "2 2 7\n")
    V, s, d, sym = G.read_matrix_market(write(tmp_path, "r.mtx", text))
    assert V == 6 and not sym
    assert s.tolist() == [0, 2, 1] and d.tolist() == [5, 0, 1]


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 2\n1 2\n", "fewer"),
    ("3 3 1\n1 2\n2 3\n", "more"),
    ("3 3 1\n4 1\n", "outside"),
    ("3 3 1\n0 1\n", "outside"),
    ("3 3 1\n1 x\n", "expected"),
    ("% only comments\n", "no size line"),
    ("%%MatrixMarket matrix array real general\n3 3\n1.0\n", "dense"),
])
def test_malformed_files_fail_loudly(tmp_path, text, msg):
    with pytest.raises(G.GraphVineError) as e:
        G.read_matrix_market(write(tmp_path, "bad.mtx", text))
    assert msg in str(e.value)


def test_missing_file(tmp_path):
    with pytest.raises(G.GraphVineError):
        G.read_matrix_market(str(tmp_path / "nope.mtx"))
