"""Whitespace-delimited matrix text, as src/io/MatrixDLM.cxx reads and writes it.

* one matrix: one row per line; the column count is the token count of the first line, the
  row count the number of (non-empty) lines, and the values are then read in row order
  regardless of line breaks (MatrixDLM.cxx:16-70);
* several matrices: a first line ``N rows cols``, a blank line, then N blocks of
  ``rows x cols`` values, each followed by a blank line (MatrixDLM.cxx:72-119, 141-166);
* writing uses the stream's default 6 significant digits (MatrixDLM.cxx:121-139).

Differences kept small on purpose: the reference counts every line read before end of
file, so a file without its final newline loses its last row there; here every non-empty
line is a row.
"""
from __future__ import annotations

import numpy as np


def _open(path):
    try:
        return open(path)
    except OSError as e:
        raise RuntimeError("error opening file") from e


def read_matrix_dlm(path):
    with _open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    if not lines:
        return np.zeros((0, 0))
    cols = len(lines[0].split())
    vals = np.array(" ".join(lines).split(), dtype=np.float64)
    rows = len(lines)
    if vals.size < rows * cols:
        raise RuntimeError(f"{path}: {rows} x {cols} values expected, {vals.size} found")
    return vals[:rows * cols].reshape(rows, cols)


def read_multiple_matrix_dlm(path):
    with _open(path) as f:
        tok = f.read().split()
    if len(tok) < 3:
        return []
    n, rows, cols = (int(float(t)) for t in tok[:3])
    vals = np.array(tok[3:3 + n * rows * cols], dtype=np.float64)
    if vals.size != n * rows * cols:
        raise RuntimeError(f"{path}: {n} matrices of {rows} x {cols} expected")
    return [m.copy() for m in vals.reshape(n, rows, cols)]


def _row(r):
    return " ".join("%g" % v for v in r)


def write_matrix_dlm(path, m):
    m = np.atleast_2d(np.asarray(m, dtype=np.float64))
    if m.size == 0:
        return
    with open(path, "w") as f:
        f.write("".join(_row(r) + "\n" for r in m))


def write_multiple_matrix_dlm(path, ms):
    ms = [np.atleast_2d(np.asarray(m, dtype=np.float64)) for m in ms]
    with open(path, "w") as f:
        if not ms:
            return
        f.write(f"{len(ms)} {ms[0].shape[0]} {ms[0].shape[1]}\n\n")
        for m in ms:
            f.write("".join(_row(r) + "\n" for r in m) + "\n")
