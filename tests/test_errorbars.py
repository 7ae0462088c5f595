"""Error-bar tool over repeated runs' TSV files (no GPU)."""
import os

import numpy as np

from laughlinedge_amd.errorbars import average_dir, parse_range


def _write(path, rows):
    with open(path, "w") as f:
        for r in rows:
            f.write("" if r is None else "\t".join(r))
            f.write("\n")


def test_matrix_occupation_and_dynamics_layouts(tmp_path):
    rng = np.random.default_rng(0)
    n = 5
    vals = rng.normal(size=(n, 3, 3, 2))
    mdir, odir, ddir = tmp_path / "Metric", tmp_path / "Occ", tmp_path / "Dyn"
    for d in (mdir, odir, ddir):
        os.makedirs(d)
    occ = rng.normal(size=(n, 2, 2, 3, 2))
    dyn = rng.normal(size=(n, 4, 2))
    for k in range(n):
        # unit diagonal as the normalized metric has: those rows agree in every file
        rows = [[str(i), str(j)] + ([f"{1.0:.6f}", f"{0.0:.6f}"] if i == j else
                                     [f"{vals[k, i, j, 0]:.6f}", f"{vals[k, i, j, 1]:.6f}"])
                for i in range(3) for j in range(3)]
        _write(mdir / f"{k}.tsv", rows)
        rows = []
        for a in range(2):
            for b in range(2):
                rows += [[str(a), str(b), str(l), f"{occ[k, a, b, l, 0]:.6f}", f"{occ[k, a, b, l, 1]:.6f}"]
                         for l in range(3)]
                rows.append(None)
        _write(odir / f"{k}.tsv", rows)
        rows = [[f"{0.0:.6f}", f"{1.0:.6f}", f"{0.0:.6f}"]] + \
               [[f"{0.5 * t:.6f}", f"{dyn[k, t, 0]:.6f}", f"{dyn[k, t, 1]:.6f}"] for t in range(1, 4)]
        _write(ddir / f"{k}.tsv", rows)
    files = parse_range("0-4")
    assert files == [0, 1, 2, 3, 4] and parse_range("0,2,5") == [0, 2, 5]
    assert average_dir(str(mdir), files, 1000)[1] == 2
    assert average_dir(str(odir), files, 1000)[1] == 3
    assert average_dir(str(ddir), files, 1000)[1] == 1
    got = [l.rstrip("\n").split("\t") for l in open(mdir / "1000.tsv")]
    err = [l.rstrip("\n").split("\t") for l in open(mdir / "1000_stderr.tsv")]
    r = vals.round(6)
    for row, erow in zip(got, err):
        i, j = int(row[0]), int(row[1])
        if i == j:
            assert float(row[2]) == 1.0 and float(erow[2]) == 0.0
            continue
        for c in range(2):
            assert abs(float(row[2 + c]) - r[:, i, j, c].mean()) < 1e-6
            assert abs(float(erow[2 + c]) - r[:, i, j, c].std(ddof=1) / np.sqrt(n)) < 1e-6
    olines = open(odir / "1000.tsv").read().split("\n")
    assert olines[3] == "" and olines[7] == ""
    first = olines[0].split("\t")
    assert first[:3] == ["0", "0", "0"] and abs(float(first[3]) - occ.round(6)[:, 0, 0, 0, 0].mean()) < 1e-6
    d0 = open(ddir / "1000.tsv").readline().split("\t")
    assert d0[0] == "0.000000" and float(d0[1]) == 1.0
