"""CPU checks of operand construction (include/segtaco/io.h).

* MatrixMarket reading against the reference's own read_mtx_file compiled
  from its sources (oracle/_ref), on the reference's 192bit.mtx fixture and on
  generated files covering every banner field / symmetry, comments, a blank
  line before the size line, duplicates and empty rows; symmetric expansion against a
  brute-force set construction.
* COO -> CSR against the reference's to_eigen_csr (Eigen setFromTriplets),
  bit-exact including the summation order of duplicates.
* The reference comparators against their compiled originals.
* Malformed files are errors, not exits.
"""
import os
import zlib

import numpy as np
import pytest

import oracle as O
import segtaco_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE_192 = "/root/reference/matrix_subset/192bit/192bit.mtx"
HAVE_REF = O.ref_lib() is not None

needs_ref = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (no reference tree)")


def _to_pos(m, rows):
    return np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))]).astype(np.int32)


def _write_mtx(path, banner, m, n, entries, comments=("% generated",), blank_before_size=False,
               values=1, rng=None):
    rng = rng or np.random.default_rng(0)
    with open(path, "w") as f:
        f.write(banner + "\n")
        for c in comments:
            f.write(c + "\n")
        if blank_before_size:  # mm_read_mtx_crd_size then scans on (mmio.h:297-306)
            f.write("\n")
        f.write(f"{m} {n} {len(entries)}\n")
        for r, c in entries:
            vals = " ".join(f"{rng.uniform(-5, 5):.6g}" for _ in range(values))
            f.write(f"{r + 1} {c + 1} {vals}".rstrip() + "\n")
    return path


@needs_ref
def test_192bit_fixture_matches_reference_reader():
    m, n, rows, cols, info = S.mtx_read(FIXTURE_192)
    rm, rn, rpos, rcrd = O.ref_read_mtx(FIXTURE_192)
    assert (m, n) == (rm, rn) == (13691, 13682)
    assert info.field == 2 and info.symmetry == 0 and info.entries == 154303
    assert np.array_equal(_to_pos(m, rows), rpos)
    assert np.array_equal(cols, rcrd)


@needs_ref
@pytest.mark.parametrize("field,values", [("pattern", 0), ("real", 1), ("integer", 1),
                                          ("complex", 2)])
@pytest.mark.parametrize("sym", ["general", "symmetric", "skew-symmetric", "hermitian"])
def test_generated_files_match_reference_reader(tmp_path, field, values, sym):
    rng = np.random.default_rng(zlib.crc32(f"{field}{sym}".encode()))
    m, n = (300, 300) if sym != "general" else (300, 250)
    k = 2000
    r = rng.integers(0, m, k)
    c = rng.integers(0, n, k)
    if sym != "general":  # lower triangle as MatrixMarket stores it
        r, c = np.maximum(r, c), np.minimum(r, c)
    r[:40] = r[40:80]  # duplicates
    c[:40] = c[40:80]
    r[r == 7] = 8  # an empty row
    entries = list(zip(r.tolist(), c.tolist()))
    banner = f"%%MatrixMarket MATRIX Coordinate {field.upper()} {sym}"
    path = _write_mtx(str(tmp_path / "a.mtx"), banner, m, n, entries,
                      comments=("% one", "%two"), blank_before_size=(sym == "hermitian"),
                      values=values, rng=rng)
    rm, rn, rpos, rcrd = O.ref_read_mtx(path)
    mm, nn, rows, cols, info = S.mtx_read(path, S.MTX_REFERENCE_TRUNCATE)
    assert (mm, nn) == (rm, rn)
    assert np.array_equal(_to_pos(mm, rows), rpos)
    assert np.array_equal(cols, rcrd)
    if sym != "symmetric":
        # truncation only concerns symmetric expansion
        _, _, rows2, cols2, _ = S.mtx_read(path)
        assert np.array_equal(rows2, rows) and np.array_equal(cols2, cols)


def test_symmetric_expansion_is_complete_by_default(tmp_path):
    rng = np.random.default_rng(3)
    m = 200
    r = rng.integers(0, m, 900)
    c = rng.integers(0, m, 900)
    r, c = np.maximum(r, c), np.minimum(r, c)
    path = _write_mtx(str(tmp_path / "s.mtx"), "%%MatrixMarket matrix coordinate pattern symmetric",
                      m, m, list(zip(r.tolist(), c.tolist())), values=0)
    _, _, rows, cols, _ = S.mtx_read(path)
    want = sorted({(a, b) for a, b in zip(r.tolist(), c.tolist())} |
                  {(b, a) for a, b in zip(r.tolist(), c.tolist())})
    assert list(zip(rows.tolist(), cols.tolist())) == want
    # the reference-compatible mode keeps only the header's entry count
    _, _, rows_t, cols_t, _ = S.mtx_read(path, S.MTX_REFERENCE_TRUNCATE)
    assert list(zip(rows_t.tolist(), cols_t.tolist())) == want[:900]


@pytest.mark.parametrize("text,what", [
    ("", "empty"),
    ("%%MatrixMarket matrix coordinate real\n3 3 1\n1 1 1\n", "banner"),
    ("%MatrixMarket matrix coordinate real general\n3 3 1\n1 1 1\n", "banner"),
    ("%%MatrixMarket vector coordinate real general\n3 3 1\n1 1 1\n", "object"),
    ("%%MatrixMarket matrix array real general\n3 3\n1\n", "array"),
    ("%%MatrixMarket matrix coordinate quaternion general\n3 3 1\n1 1 1\n", "field"),
    ("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1\n", "short"),
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1\n", "range"),
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n0 1 1\n", "range"),
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 1\n", "value"),
    ("%%MatrixMarket matrix coordinate real general\n% only comments\n", "size"),
    # a non-square "symmetric" file whose mirrored entry (2, 5) leaves 5 x 3
    ("%%MatrixMarket matrix coordinate real symmetric\n5 3 1\n5 2 1\n", "mirror"),
])
def test_malformed_files_are_errors(tmp_path, text, what):
    path = tmp_path / "bad.mtx"
    path.write_text(text)
    with pytest.raises(S.SegtacoError):
        S.mtx_read(str(path))


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(S.SegtacoError, match="not found"):
        S.mtx_read(str(tmp_path / "nope.mtx"))


@needs_ref
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_coo_to_csr_matches_eigen_set_from_triplets(seed):
    rng = np.random.default_rng(seed)
    m, n, k = 500, 400, 6000
    rows = rng.integers(0, m, k).astype(np.int32)
    cols = rng.integers(0, n, k).astype(np.int32)
    # many duplicates, values whose sum depends on the order of addition
    rows[:3000] = rows[3000:]
    cols[:3000:2] = cols[3000::2]
    vals = rng.choice(np.array([1e8, 1.0, -1e8, 0.1, 0.0, 3.5e-3], np.float32), k)
    rows[rows == 11] = 12  # empty row
    pos, crd, v = S.coo_to_csr(m, n, rows, cols, vals)
    rpos, rcrd, rv = O.ref_coo_to_csr(m, n, rows, cols, vals)
    assert np.array_equal(pos, rpos) and np.array_equal(crd, rcrd)
    assert np.array_equal(v.view(np.uint32), rv.view(np.uint32))
    assert np.any(v == 0.0)  # explicit zeros are kept


def test_coo_to_csr_pattern_and_errors():
    pos, crd, v = S.coo_to_csr(3, 3, [2, 0, 2, 0], [1, 2, 1, 0])
    assert pos.tolist() == [0, 2, 2, 3] and crd.tolist() == [0, 2, 1] and v is None
    with pytest.raises(S.SegtacoError):
        S.coo_to_csr(3, 3, [3], [0], [1.0])
    with pytest.raises(S.SegtacoError):
        S.coo_to_csr(3, 3, [0], [-1], [1.0])


def test_reference_value_fill():
    a = S.fill_reference_vals(7, 30000)
    assert np.array_equal(a, S.fill_reference_vals(7, 30000))
    assert not np.array_equal(a, S.fill_reference_vals(8, 30000))
    assert set(np.unique(a).tolist()) == {np.float32(0.0), np.float32(0.1), np.float32(0.2)}
    assert np.all(a[a > 0.15] == np.float32(2) / np.float32(10))
    counts = np.bincount(np.rint(a * 10).astype(int), minlength=3)
    assert np.all(np.abs(counts / 30000 - 1 / 3) < 0.02)


def test_read_csr_is_read_fill_canonicalise():
    path = FIXTURE_192 if os.path.exists(FIXTURE_192) else None
    if path is None:
        pytest.skip("reference fixture absent")
    M, K, pos, crd, vals = S.mtx_read_csr(path, seed=11)
    m, n, rows, cols, _ = S.mtx_read(path)
    p2, c2, v2 = S.coo_to_csr(m, n, rows, cols, S.fill_reference_vals(11, len(rows)))
    assert (M, K) == (m, n)
    assert np.array_equal(pos, p2) and np.array_equal(crd, c2) and np.array_equal(vals, v2)
    _, _, _, _, ones = S.mtx_read_csr(path, seed=2**64 - 1)
    assert np.all(ones == 1.0)


@needs_ref
def test_comparators_match_reference():
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, 5000).astype(np.float32)
    y = x.copy()
    y[::97] += rng.uniform(-0.01, 0.01, len(y[::97])).astype(np.float32)
    assert S.compare_array(x, y) == O.ref_compare_array(x, y)
    assert S.compare_array_my(x, y) == O.ref_compare_array_my(x, y)
    # x == 0 with y != 0 -> infinity; x == y == 0 ignored
    x0 = np.array([0.0, 1.0, 0.0], np.float32)
    assert S.compare_array(x0, np.array([0.0, 1.5, 0.0], np.float32)) == pytest.approx(0.5)
    assert S.compare_array(x0, np.array([0.0, 1.0, 1e-30], np.float32)) == np.inf
    assert O.ref_compare_array(x0, np.array([0.0, 1.0, 1e-30], np.float32)) == np.inf
    # the mismatch count stops once it passes 10000
    big = np.zeros(30000, np.float32)
    assert S.compare_array_my(big, big + 1) == 10001.0
    assert S.compare_matrices(big.reshape(100, 300), big.reshape(100, 300)) == 0.0
    assert S.compare_matrices(big.reshape(100, 300), big.reshape(300, 100)) == np.inf
