"""On-disk interchange of assembled SHE systems (MatrixMarket).

The reference includes ViennaCL's MatrixMarket reader and writer next to its
solvers (src/solvers/viennacl/serial_linear_solver.hpp:31) but never calls
them.  Here they carry systems between the GPU fill, the oracle and test
fixtures.  A system with prefix P is stored as:
  P.mtx           the matrix: coordinate real general, 1-based, rows in order,
                  columns ascending within a row, explicit zeros kept;
  P_b.mtx         the right-hand side: array real general, n x 1;
  P_blocks.mtx    optional block ranges (fill_block_indices pairs):
                  array integer general, nblocks x 2, 0-based [first, last);
  P_mapping.mtx   optional even mapping: array integer general, (cells * N_H) x 2,
                  first unknown index (0-based, -1 = none) and expansion order
                  per (cell, H), cell-major.
Values are written with 17 significant digits, so FP64 round-trips exactly.
"""
from __future__ import annotations

import numpy as np


def _header(f, kind, field):
    f.write(f"%%MatrixMarket matrix {kind} {field} general\n".encode())


def _read_lines(path):
    with open(path, "rb") as f:
        head = f.readline().decode().strip().split()
        if len(head) < 5 or head[0] != "%%MatrixMarket" or head[1] != "matrix":
            raise ValueError(f"{path}: not a MatrixMarket matrix file")
        kind, field, sym = head[2].lower(), head[3].lower(), head[4].lower()
        if sym != "general":
            raise ValueError(f"{path}: only general matrices are supported")
        line = f.readline()
        while line.startswith(b"%"):
            line = f.readline()
        dims = [int(t) for t in line.split()]
        body = f.read()
    return kind, field, dims, body


def write_matrix(path, row_ptr, col, val):
    row_ptr = np.asarray(row_ptr, np.int64)
    col = np.asarray(col, np.int64)
    val = np.asarray(val, np.float64)
    n = len(row_ptr) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(row_ptr))
    with open(path, "wb") as f:
        _header(f, "coordinate", "real")
        f.write(f"{n} {n} {len(val)}\n".encode())
        if len(val):
            np.savetxt(f, np.column_stack([rows + 1, col + 1, val]), fmt=["%d", "%d", "%.17g"])


def read_matrix(path):
    """CSR (row_ptr int64, col int32, val float64) with entries sorted by (row, col)."""
    kind, field, dims, body = _read_lines(path)
    if kind != "coordinate" or field not in ("real", "integer"):
        raise ValueError(f"{path}: expected a coordinate real matrix")
    n, m, nnz = dims
    if n != m:
        raise ValueError(f"{path}: matrix is not square")
    data = np.loadtxt(body.decode().splitlines(), ndmin=2) if nnz else np.zeros((0, 3))
    if data.shape[0] != nnz:
        raise ValueError(f"{path}: {data.shape[0]} entries, header says {nnz}")
    r = data[:, 0].astype(np.int64) - 1
    c = data[:, 1].astype(np.int64) - 1
    v = data[:, 2].astype(np.float64)
    if nnz and (r.min() < 0 or r.max() >= n or c.min() < 0 or c.max() >= n):
        raise ValueError(f"{path}: index out of range")
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    row_ptr = np.zeros(n + 1, np.int64)
    np.add.at(row_ptr, r + 1, 1)
    return np.cumsum(row_ptr), c.astype(np.int32), v


def write_array(path, a, integer=False):
    a = np.asarray(a)
    a2 = a.reshape(a.shape[0], -1)
    with open(path, "wb") as f:
        _header(f, "array", "integer" if integer else "real")
        f.write(f"{a2.shape[0]} {a2.shape[1]}\n".encode())
        # MatrixMarket arrays are column-major
        np.savetxt(f, a2.T.reshape(-1), fmt="%d" if integer else "%.17g")


def read_array(path):
    kind, field, dims, body = _read_lines(path)
    if kind != "array":
        raise ValueError(f"{path}: expected an array")
    rows, cols = dims
    vals = np.array(body.split(), dtype=np.float64 if field == "real" else np.int64)
    if vals.size != rows * cols:
        raise ValueError(f"{path}: {vals.size} values, header says {rows * cols}")
    return vals.reshape(cols, rows).T


def write_system(prefix, row_ptr, col, val, b, block_ranges=None, mapping=None):
    write_matrix(f"{prefix}.mtx", row_ptr, col, val)
    write_array(f"{prefix}_b.mtx", np.asarray(b, np.float64))
    if block_ranges is not None:
        write_array(f"{prefix}_blocks.mtx", np.asarray(block_ranges, np.int64).reshape(-1, 2), integer=True)
    if mapping is not None:
        idx, order = mapping
        write_array(f"{prefix}_mapping.mtx",
                    np.column_stack([np.asarray(idx, np.int64).ravel(), np.asarray(order, np.int64).ravel()]),
                    integer=True)


def read_system(prefix, blocks=False, mapping=False):
    """Returns (row_ptr, col, val, b[, block_ranges][, (idx, order)])."""
    rp, ci, v = read_matrix(f"{prefix}.mtx")
    b = read_array(f"{prefix}_b.mtx")[:, 0]
    if len(b) != len(rp) - 1:
        raise ValueError(f"{prefix}: right-hand side has {len(b)} entries for {len(rp) - 1} rows")
    out = [rp, ci, v, b]
    if blocks:
        out.append(read_array(f"{prefix}_blocks.mtx").astype(np.int64))
    if mapping:
        m = read_array(f"{prefix}_mapping.mtx").astype(np.int64)
        out.append((m[:, 0], m[:, 1].astype(np.int32)))
    return tuple(out)
