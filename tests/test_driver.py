"""The sweep driver (segtaco_amd/matrix-eval-gpu).

CPU: argument handling and input errors (they end before any device call),
and that the driver sees only the C-ABI headers. GPU: a full 61-row sweep
writes the reference's CSV rows in the reference's order, --summary writes
the per-matrix minimum table, and the verify experiment finds every
operator-surface entry point equal to the vendor-library row.
"""
import csv
import os
import re
import subprocess

import numpy as np
import pytest

import segtaco_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "segtaco_amd", "matrix-eval-gpu")
SRC = os.path.join(ROOT, "segtaco_amd", "driver", "matrix_main.cpp")
GOLD = os.path.join(ROOT, "tests", "golden", "mtx192bit_head1500_n2.npz")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")

SWEEP_LABELS = ([("rb", "pr", f"tune{i}") for i in range(30)] +
                [("eb", "pr", f"tune{i}") for i in range(12)] +
                [("rb", "sr", f"tune{i}") for i in range(9)] +
                [("eb", "sr", f"tune{i}") for i in range(9)] + [("eb", "sr", "alg2")])


def _run(args, timeout=300):
    return subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=timeout)


def _write_mtx_from_golden(path, dup_lines=25):
    """The 1500-row head of the reference's 192bit.mtx (pattern, general),
    written back as MatrixMarket with shuffled lines and some duplicates."""
    g = np.load(GOLD)
    pos, crd = g["pos"], g["crd"]
    M, K = int(g["M"]), int(g["K"])
    rows = np.repeat(np.arange(M), np.diff(pos))
    rng = np.random.default_rng(4)
    order = rng.permutation(len(crd))
    extra = rng.integers(0, len(crd), dup_lines)
    lines = [f"{rows[i] + 1} {crd[i] + 1}" for i in np.concatenate([order, extra])]
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern general\n% head of 192bit\n")
        f.write(f"{M} {K} {len(lines)}\n" + "\n".join(lines) + "\n")
    return M, K, len(crd)


def test_driver_includes_no_hip_header():
    text = open(SRC).read()
    includes = re.findall(r'#include\s*[<"]([^>"]+)[>"]', text)
    assert all(not i.startswith("hip") for i in includes)
    assert {"segtaco/spmm.h", "segtaco/io.h"} <= set(includes)


def test_bad_arguments_exit_2():
    r = _run(["d", "n", "spmm_csr_gpu", ".", "hw", "--cols", "0"])
    assert r.returncode == 2 and "usage" in r.stderr
    r = _run(["d", "n", "spmm_csr_gpu", ".", "hw", "--bogus"])
    assert r.returncode == 2 and "unknown option" in r.stderr
    r = _run(["d", "n", "spmm_csr_gpu", ".", "hw", "--synthetic", "cube:10:2"])
    assert r.returncode == 2


def test_missing_matrix_is_an_error(tmp_path):
    r = _run([str(tmp_path), "nope", "spmm_csr_gpu", str(tmp_path), "hw"])
    assert r.returncode == 1 and "not found" in r.stderr
    assert not list(tmp_path.iterdir())  # no CSV opened before the input loaded


@pytest.mark.gpu
def test_sweep_writes_reference_rows(tmp_path):
    _write_mtx_from_golden(tmp_path / "head.mtx")
    r = _run([str(tmp_path), "head", "spmm_csr_gpu", str(tmp_path), "mi355x", "--trials", "3",
              "--cols", "8", "--summary"])
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "nnz: 16355 K: 8" in r.stdout
    with open(tmp_path / "spmm_csr_gpu_8_mi355x.csv") as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["kernel", "platform", "format", "library", "tensor", "time"]
    body = rows[1:]
    assert len(body) == 61 * 3
    for i, (kern, plat, fmt, lib, tensor, ms) in enumerate(body):
        assert (kern, tensor) == ("spmm", "head")
        assert (plat, fmt, lib) == SWEEP_LABELS[i // 3]
        assert float(ms) > 0
    with open(tmp_path / "spmm_csr_gpu_8_mi355x_summary.csv") as f:
        summ = list(csv.reader(f))
    assert summ[0][0] == "dataset" and summ[0][1] == "rbprtune0" and summ[0][-1] == "alg2"
    assert len(summ[0]) == 62 and summ[1][0] == "head"
    mins = [min(float(body[3 * r + t][5]) for t in range(3)) for r in range(61)]
    assert np.allclose([float(x) for x in summ[1][1:]], mins, rtol=1e-5)
    # resana's other table: the mean over the trials, rounded to 4 decimals
    with open(tmp_path / "spmm_csr_gpu_8_mi355x_summary_mean.csv") as f:
        summ_m = list(csv.reader(f))
    assert summ_m[0] == summ[0] and summ_m[1][0] == "head"
    means = [round(sum(float(body[3 * r + t][5]) for t in range(3)) / 3, 4) for r in range(61)]
    assert np.allclose([float(x) for x in summ_m[1][1:]], means, atol=1.5e-4)
    # a second run appends without a second header
    r = _run([str(tmp_path), "head", "spmm_csr_gpu", str(tmp_path), "mi355x", "--trials", "1",
              "--cols", "8"])
    assert r.returncode == 0
    with open(tmp_path / "spmm_csr_gpu_8_mi355x.csv") as f:
        rows2 = list(csv.reader(f))
    assert len(rows2) == 1 + 61 * 4 and rows2.count(rows[0]) == 1


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [["--random-b", "--cols", "33"],
                                   ["--synthetic", "powerlaw:20000:12", "--random-b"]])
def test_verify_every_entry_point_against_vendor_row(tmp_path, extra):
    _write_mtx_from_golden(tmp_path / "head.mtx")
    r = _run([str(tmp_path), "head", "spmm_csr_gpu_verify", str(tmp_path), "hw"] + extra)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    lines = [l for l in r.stdout.splitlines() if "vs rocsparse" in l]
    assert len(lines) == S.lib().segtaco_legacy_count()
    assert all(l.endswith("vs rocsparse: 0") for l in lines), lines
    assert r.returncode == 0
