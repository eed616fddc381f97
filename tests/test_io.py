"""File formats of the reference: data matrices, k-NN graph text, DPG layers.

Files are written here byte by byte from the reference's record layouts
(src/iodelegator.cpp:186-385, src/abstractmerge.cpp:278-291,
src/lynndescent.cpp:625-638, src/lysearch.cpp:112-186) and read through the
library, and the library's writers are read back.  Host code only: no GPU.
"""
import struct

import numpy as np
import pytest

import nnm


def test_fvecs_records(tmp_path):
    rng = np.random.default_rng(1)
    X = rng.standard_normal((37, 12)).astype(np.float32)
    p = tmp_path / "a.fvecs"
    with open(p, "wb") as f:
        for row in X:
            f.write(struct.pack("<i", 12))
            f.write(row.tobytes())
    assert np.array_equal(nnm.read_matrix(str(p)), X)
    q = tmp_path / "b.fvecs"
    nnm.write_matrix(str(q), X)
    assert q.read_bytes() == p.read_bytes()


def test_bvecs_records_and_empty_rows(tmp_path):
    rng = np.random.default_rng(2)
    B = rng.integers(0, 256, size=(20, 16), dtype=np.uint8)
    p = tmp_path / "a.bvecs"
    with open(p, "wb") as f:
        for i, row in enumerate(B):
            f.write(struct.pack("<i", 16))
            f.write(row.tobytes())
            if i == 4:
                f.write(struct.pack("<i", 0))  # empty record: skipped by the reader
    X = nnm.read_matrix(str(p), max_rows=20)
    assert X.dtype == np.float32 and np.array_equal(X, B.astype(np.float32))
    q = tmp_path / "b.bvecs"
    nnm.write_matrix(str(q), B.astype(np.float32))
    assert np.array_equal(nnm.read_matrix(str(q)), B.astype(np.float32))
    with pytest.raises(nnm.NNMError):
        nnm.write_matrix(str(tmp_path / "c.bvecs"), np.array([[0.5, 3.0]], np.float32))  # not a byte value


def test_text_matrix(tmp_path):
    p = tmp_path / "m.txt"
    p.write_text("3 4\n1 2 3 4\n5 6 7 8\n9 10 11 12.5\n")
    X = nnm.read_matrix(str(p))
    assert X.shape == (3, 4) and X[2, 3] == np.float32(12.5) and X[1, 0] == 5
    q = tmp_path / "n.txt"
    Y = np.random.default_rng(3).standard_normal((5, 7)).astype(np.float32)
    nnm.write_matrix(str(q), Y)
    assert np.array_equal(nnm.read_matrix(str(q)), Y)  # %.9g round-trips float32


def test_knn_graph_text(tmp_path):
    p = tmp_path / "g.txt"
    # rows carry 5 ids; a reader asked for topk = 3 keeps the first 3
    p.write_text("3 5\n0 5 1 2 3 4 5 \n1 5 0 2 3 4 5 \n2 5 9 8 7 6 5 \n")
    ids = nnm.read_knn(str(p), topk=3)
    assert ids.tolist() == [[1, 2, 3], [0, 2, 3], [9, 8, 7]]
    g = np.random.default_rng(4).integers(0, 1000, size=(50, 20)).astype(np.uint32)
    q = tmp_path / "h.txt"
    nnm.write_knn(str(q), g)
    first = q.read_text().splitlines()[:2]
    assert first[0] == "50 20" and first[1].split()[:3] == ["0", "20", str(g[0, 0])]
    assert np.array_equal(nnm.read_knn(str(q)), g)


def test_layered_dpg_text(tmp_path):
    p = tmp_path / "l.txt"
    # two blocks, upper layer (3 rows) first, then the base layer (5 rows)
    p.write_text("3\n0 2 1 2 \n1 1 0 \n2 1 0 \n5\n0 2 1 4 \n1 1 0 \n2 0 \n3 1 2 \n4 2 3 0 \n")
    layers = nnm.read_layers(str(p))
    assert len(layers) == 2
    (o0, a0), (o1, a1) = layers
    assert o0.tolist() == [0, 2, 3, 4] and a0.tolist() == [1, 2, 0, 0]
    assert o1.tolist() == [0, 2, 3, 3, 4, 6] and a1.tolist() == [1, 4, 0, 2, 3, 0]
    q = tmp_path / "m.txt"
    nnm.write_layers(str(q), layers)
    again = nnm.read_layers(str(q))
    for (x, y), (u, w) in zip(layers, again):
        assert np.array_equal(x, u) and np.array_equal(y, w)


def test_malformed_files_fail(tmp_path):
    with pytest.raises(nnm.NNMError):
        nnm.read_matrix(str(tmp_path / "missing.fvecs"))
    p = tmp_path / "bad.txt"
    p.write_text("2 3\n1 2 3\n4 5\n")  # short
    with pytest.raises(nnm.NNMError):
        nnm.read_matrix(str(p))
    with pytest.raises(nnm.NNMError):
        nnm.read_matrix(str(tmp_path / "x.unknown"))


def test_reference_sift_fixture_reads():
    """The first 1000 rows of the reference's data/sift_query.txt (text
    matrix): integers 0..255, parsed the same as a plain split."""
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sift_query_head1000.txt")
    X = nnm.read_matrix(p)
    assert X.shape == (1000, 128)
    with open(p) as f:
        f.readline()
        first = np.array(f.readline().split(), np.float32)
    assert np.array_equal(X[0], first)
    assert X.min() >= 0 and X.max() <= 255 and np.array_equal(X, np.rint(X))
