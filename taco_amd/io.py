"""Matrix CSV files in the reference's layout (io.hpp:14-38): a header line
"row,col,re,im", then one line per entry in row-major order with the row and
column indices and the real and imaginary parts in fixed notation with 17
digits after the decimal point.  The driver writes GW, L4 and the final
density matrix this way."""
from __future__ import annotations

import os
from typing import IO, Union

import numpy as np

__all__ = ["write_csv_matrix", "read_csv_matrix"]

PathOrFile = Union[str, os.PathLike, IO[str]]


def write_csv_matrix(dest: PathOrFile, M, precision: int = 17) -> None:
    """Writes the complex matrix M (io.hpp:14-38)."""
    M = np.asarray(M, np.complex128)
    if M.ndim != 2:
        raise ValueError("write_csv_matrix: M must be a matrix")
    fmt = f"{{}},{{}},{{:.{precision}f}},{{:.{precision}f}}\n"
    lines = ["row,col,re,im\n"]
    for r in range(M.shape[0]):
        for c in range(M.shape[1]):
            v = M[r, c]
            lines.append(fmt.format(r, c, float(v.real), float(v.imag)))
    text = "".join(lines)
    if hasattr(dest, "write"):
        dest.write(text)
        return
    try:
        with open(dest, "w") as f:
            f.write(text)
    except OSError as e:
        raise RuntimeError(f"Failed to open output file: {os.fspath(dest)}") from e


def read_csv_matrix(src: PathOrFile) -> np.ndarray:
    """Reads a matrix written by write_csv_matrix (or the reference's
    write_csv_matrix); the shape is one more than the largest indices."""
    text = src.read() if hasattr(src, "read") else open(src).read()
    lines = [ln for ln in text.splitlines() if ln.strip()]
    if not lines or lines[0].strip() != "row,col,re,im":
        raise ValueError("read_csv_matrix: missing header 'row,col,re,im'")
    rows = np.array([ln.split(",") for ln in lines[1:]], dtype=object).reshape(-1, 4)
    r = rows[:, 0].astype(np.int64)
    c = rows[:, 1].astype(np.int64)
    nr = int(r.max()) + 1 if r.size else 0
    nc = int(c.max()) + 1 if c.size else 0
    M = np.zeros((nr, nc), np.complex128)
    M[r, c] = rows[:, 2].astype(np.float64) + 1j * rows[:, 3].astype(np.float64)
    return M
