"""Error bars over repeated runs: the TSV files that the drop-in interface writes for several
fileNumbers (independent seeds) are averaged entry by entry.

The reference leaves this step to external scripts (its main.cu repeats a computation with
fileNumber = 0, 1, ... and writes one file per repetition, modules/edge_mc/em_spectrum.cu:393-414,
em_edgeResponse.tpp:559-594, gf_occupation_numbers.cu:515-538, dynamics.cpp:293-309).  Every
file of one directory has the same rows: leading key columns (indices, or the time of the
dynamics files) and trailing value columns, with blank lines between blocks in the
occupation files.  The tool writes

    DIR/<out>.tsv         the means, in the input layout ("%.6f" values), readable by
                          timeEvolution::edgeDensityResponse_compute as fileNumber <out>
    DIR/<out>_stderr.tsv  the standard errors of the means (sample standard deviation over
                          the files / sqrt(number of files))

Usage: python -m laughlinedge_amd.errorbars DIR [DIR ...] --files 0-9 [--out 1000]
"""
import argparse
import math
import os
import sys


def parse_range(spec):
    """'0-9' or '0,2,5' or '3' -> list of ints"""
    out = []
    for part in spec.split(","):
        part = part.strip()
        if "-" in part:
            a, b = part.split("-", 1)
            out.extend(range(int(a), int(b) + 1))
        elif part:
            out.append(int(part))
    return out


def _read(path):
    rows = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            rows.append(line.split("\t") if line.strip() else None)
    return rows


def average_dir(directory, files, out):
    """Average DIR/{k}.tsv over k in files; returns (number of rows, number of key columns)."""
    if len(files) < 2:
        raise ValueError("need at least two files for error bars")
    tables = [_read(os.path.join(directory, f"{k}.tsv")) for k in files]
    n_rows = len(tables[0])
    if any(len(t) != n_rows for t in tables):
        raise ValueError(f"{directory}: the files have different numbers of rows")
    # key columns: the leading columns equal in every file on every row (a row whose values
    # happen to agree, such as a unit diagonal or the start of a time series, does not count)
    width = None
    for r in range(n_rows):
        cells = [t[r] for t in tables]
        if cells[0] is None:
            if any(c is not None for c in cells):
                raise ValueError(f"{directory}: blank line {r + 1} is not blank in every file")
            continue
        if any(c is None or len(c) != len(cells[0]) for c in cells):
            raise ValueError(f"{directory}: row {r + 1} differs in shape between files")
        nkey = 0
        while nkey < len(cells[0]) and all(c[nkey] == cells[0][nkey] for c in cells):
            nkey += 1
        width = nkey if width is None else min(width, nkey)
    width = width or 0
    mean_lines, err_lines = [], []
    for r in range(n_rows):
        cells = [t[r] for t in tables]
        if cells[0] is None:
            mean_lines.append("")
            err_lines.append("")
            continue
        keys = cells[0][:width]
        means, errs = [], []
        for col in range(width, len(cells[0])):
            v = [float(c[col]) for c in cells]
            m = sum(v) / len(v)
            var = sum((x - m) ** 2 for x in v) / (len(v) - 1)
            means.append(m)
            errs.append(math.sqrt(var / len(v)))
        mean_lines.append("\t".join(keys + [f"{x:.6f}" for x in means]))
        err_lines.append("\t".join(keys + [f"{x:.6f}" for x in errs]))
    with open(os.path.join(directory, f"{out}.tsv"), "w") as f:
        f.write("\n".join(mean_lines) + "\n")
    with open(os.path.join(directory, f"{out}_stderr.tsv"), "w") as f:
        f.write("\n".join(err_lines) + "\n")
    return n_rows, width


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("dirs", nargs="+", help="directories holding <fileNumber>.tsv files")
    ap.add_argument("--files", required=True, help="file numbers, e.g. 0-9 or 0,1,4")
    ap.add_argument("--out", type=int, default=1000, help="file number of the outputs")
    args = ap.parse_args(argv)
    files = parse_range(args.files)
    for d in args.dirs:
        rows, _ = average_dir(d, files, args.out)
        print(f"{d}: {len(files)} files, {rows} rows -> {args.out}.tsv, {args.out}_stderr.tsv")
    return 0


if __name__ == "__main__":
    sys.exit(main())
