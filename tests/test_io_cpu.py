"""Host-side ingest of the reference's file formats (no GPU): binary pairs
(tc/graph/graph.cpp:23-39), text edge lists (Txt2Bin, graph.cpp:179-214) and Matrix Market
coordinate files (bfs/graph/graph.cpp:17-122), on hand-written fixtures and error cases."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_oracle, load_rtg

rtg = load_rtg()


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    return p


def test_binary_pairs_round_trip_and_odd_sizes(tmp_path):
    pairs = np.array([1, 2, 3, 4, -1, 7], np.int32)
    p = tmp_path / "g.bin"
    pairs.astype("<i4").tofile(p)
    assert np.array_equal(rtg.read_binary_pairs(p), pairs)
    # a trailing partial pair is ignored: edges = file_size / 8
    write(tmp_path, "odd.bin", pairs.astype("<i4").tobytes() + b"\x01\x02\x03")
    assert np.array_equal(rtg.read_binary_pairs(tmp_path / "odd.bin"), pairs)
    assert rtg.read_binary_pairs(write(tmp_path, "e.bin", b"")).size == 0


def test_text_pairs_like_txt2bin(tmp_path):
    p = write(tmp_path, "g.txt", "1 2\n# comment\ngarbage\n3\n4 5 6\n\n99999999999 1\n-5 7\n 8\t9\n")
    # lines that do not begin with two int32 values are skipped; extra columns are ignored
    assert rtg.read_text_pairs(p).tolist() == [1, 2, 4, 5, -5, 7, 8, 9]
    assert rtg.read_text_pairs(write(tmp_path, "e.txt", "")).size == 0
    out = tmp_path / "g.bin"
    rtg.Graph.Txt2Bin(p, out)
    assert rtg.read_binary_pairs(out).tolist() == [1, 2, 4, 5, -5, 7, 8, 9]


def test_text_pairs_int_extraction_edge_cases(tmp_path):
    """Each line is read as `stream >> int32 >> int32` reads it: leading blanks and CR skipped,
    an optional sign, digits up to the first other character, int32 bounds, and anything after
    the second value ignored; the last line needs no newline."""
    text = ("1 2\r\n+5 -6\n12abc 5\n5 8.9\n0x10 5\n  \t \n2147483647 -2147483648\n"
            "2147483648 0\n-2147483649 0\n7\t\t8 junk\n-0 +0\n- 1\n9 10")
    got = rtg.read_text_pairs(write(tmp_path, "edge.txt", text)).tolist()
    assert got == [1, 2, 5, -6, 5, 8, 2147483647, -2147483648, 7, 8, 0, 0, 9, 10]


def test_text_and_mtx_readers_at_scale(tmp_path):
    rng = np.random.default_rng(4)
    pairs = rng.integers(0, 1 << 20, size=(200000, 2)).astype(np.int32)
    pt = tmp_path / "g.txt"
    np.savetxt(pt, pairs, fmt="%d")
    assert np.array_equal(rtg.read_text_pairs(pt), pairs.reshape(-1))
    pm = tmp_path / "g.mtx"
    with open(pm, "w") as f:
        f.write("%%MatrixMarket matrix coordinate integer general\n% generated\n")
        f.write(f"{1 << 20} {1 << 20} {pairs.shape[0]}\n")
        np.savetxt(f, np.column_stack([pairs + 1, rng.integers(-5, 5, size=pairs.shape[0])]), fmt="%d")
    assert np.array_equal(rtg.read_mtx_pairs(pm), pairs.reshape(-1))


def test_mtx_pattern_symmetric_expands_off_diagonal(tmp_path):
    p = write(tmp_path, "s.mtx", "%%MatrixMarket matrix coordinate pattern symmetric\n% c\n3 3 3\n1 2\n2 3\n3 3\n")
    # (r, c) then (c, r) for off-diagonal entries, diagonal once, 1-based -> 0-based
    assert rtg.read_mtx_pairs(p).tolist() == [0, 1, 1, 0, 1, 2, 2, 1, 2, 2]


def test_mtx_real_general_reads_nnz_entries(tmp_path):
    p = write(tmp_path, "r.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 2\n1 2 0.5\n3 1 -2\n2 2 9\n")
    assert rtg.read_mtx_pairs(p).tolist() == [0, 1, 2, 0]  # the header's nnz = 2 entries


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix coordinate real general\n3 3 2\n1 2 0.5\n",          # fewer entries than nnz
    "%%MatrixMarket matrix coordinate integer general\n2 2 1\n0 1 1\n",         # zero-indexed
    "%%MatrixMarket matrix coordinate pattern general\n2 2 1\n3 1\n",           # outside the matrix
    "%%MatrixMarket matrix coordinate pattern general\n2 2 1\n" + "1" * 3000 + " 2\n",
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",              # dense array
    "%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1 1\n",       # unsupported field
    "%%MatrixMarket matrix coordinate pattern general\n",                       # no size line
    "not a banner\n1 1 1\n1 1\n",
], ids=["short", "zero-indexed", "outside", "overflow", "array", "complex", "no-size", "no-banner"])
def test_mtx_errors(tmp_path, text):
    with pytest.raises(rtg.RTGError) as e:
        rtg.read_mtx_pairs(write(tmp_path, "bad.mtx", text))
    assert e.value.code == 4  # RTG_EIO


def test_missing_files(tmp_path):
    for fn in (rtg.read_text_pairs, rtg.read_mtx_pairs):
        with pytest.raises(rtg.RTGError):
            fn(tmp_path / "missing")


def test_csr_cache_file_round_trip_and_rejections(tmp_path):
    """The preprocessed-CSR cache: exact round trip, and refusal of another key, a flipped
    byte, a truncated file and a file of another kind."""
    orc = load_oracle()
    pairs = rtg.gen_ba(2000, 4, seed=3)
    off, adj = orc.csr_from_pairs(pairs)
    key = rtg.pairs_fingerprint(pairs)
    assert key == rtg.pairs_fingerprint(pairs.copy()) and key != rtg.pairs_fingerprint(pairs, flags=1)
    p = tmp_path / "g.csr"
    rtg.csr_file_write(p, off, adj, key=key)
    o2, a2, oriented = rtg.csr_file_read(p, key=key)
    assert oriented and np.array_equal(o2, off) and np.array_equal(a2, adj)
    assert not (tmp_path / "g.csr.part").exists()
    with pytest.raises(rtg.RTGError):
        rtg.csr_file_read(p, key=key + 1)
    raw = bytearray(p.read_bytes())
    raw[100] ^= 1
    (tmp_path / "flip.csr").write_bytes(bytes(raw))
    with pytest.raises(rtg.RTGError):
        rtg.csr_file_read(tmp_path / "flip.csr", key=key)
    (tmp_path / "short.csr").write_bytes(p.read_bytes()[:-4])
    with pytest.raises(rtg.RTGError):
        rtg.csr_file_read(tmp_path / "short.csr", key=key)
    with pytest.raises(rtg.RTGError):
        rtg.csr_file_read(write(tmp_path, "x.csr", "%%MatrixMarket matrix coordinate pattern general\n"), key=key)
    # an empty graph and a symmetric (unoriented) one
    rtg.csr_file_write(tmp_path / "e.csr", np.zeros(1, np.int32), np.zeros(0, np.int32), key=7, oriented=False)
    o3, a3, or3 = rtg.csr_file_read(tmp_path / "e.csr", key=7)
    assert o3.tolist() == [0] and a3.size == 0 and not or3

