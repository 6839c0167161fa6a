"""Independent reader of the library's on-disk ELL format (.qell), test
infrastructure only: it parses the file with numpy alone (no libqmb) and
returns the matrix in the reference's row-major layout (ind[n*W], val[n*W],
track[n]), the layout of BuildHamiltonianXXZForAbelianGroup
(QMBCuda/exact_diagonalization/utils.cu:83-184), so that a saved matrix can be
compared bit for bit with the oracle's own build.

Format (little-endian): a 104-byte header, then chunk_ptr (int64, n_chunks+1),
row_len (uint16, n_rows), col (uint32, stored), then either val (float64,
stored x 1 or 2) or, packed, diag (float64, n_rows x 1 or 2), the value table
(float64, n_values) and, complex, the imaginary-part table (float64,
n_values_im). Rows sit in chunks of 64, slot-major: slot s of row r is at
chunk_ptr[r // 64] + 64 s + r % 64. A packed off-diagonal word is
col | (index << col_bits) (complex: col | (re << col_bits) | (im << (col_bits +
re_bits))); slot 0 of a packed row is its diagonal.
"""
from __future__ import annotations

import numpy as np

HEADER = np.dtype([
    ("magic", "S8"), ("version", "<u4"), ("header_bytes", "<u4"),
    ("value_type", "<i4"), ("packed", "<i4"), ("col_bits", "<i4"), ("n_values", "<i4"), ("max_terms", "<i4"),
    ("max_row_len", "<i4"),
    ("n_rows", "<i8"), ("n_chunks", "<i8"), ("nnz", "<i8"), ("stored", "<i8"),
    ("n_sites", "<i4"), ("n_up", "<i4"), ("irrep", "<i4"), ("layout", "<i4"),
    ("group_hash", "<u8"), ("re_bits", "<i4"), ("n_values_im", "<i4"),
])
assert HEADER.itemsize == 104
MAGIC = b"QMBELL01"


def _cplx(re, im):
    out = np.empty(len(re), np.complex128)
    out.real = re
    out.imag = im
    return out


def read_qell(path):
    """Parse a .qell file. Returns (header dict, ind, val, track) in the
    reference layout (val complex128 for a complex matrix)."""
    raw = np.fromfile(path, np.uint8)
    h = np.frombuffer(raw[:HEADER.itemsize].tobytes(), HEADER)[0]
    if h["magic"] != MAGIC or h["version"] != 2 or h["header_bytes"] != HEADER.itemsize:
        raise ValueError("not a version-2 ELL file")
    hd = {k: (h[k].item() if k != "magic" else bytes(h[k])) for k in HEADER.names}
    n, nc, stored, W = hd["n_rows"], hd["n_chunks"], hd["stored"], hd["max_terms"]
    cplx = hd["value_type"] == 1
    vs = 2 if cplx else 1
    off = HEADER.itemsize

    def take(dtype, count):
        nonlocal off
        nb = np.dtype(dtype).itemsize * count
        if off + nb > len(raw):
            raise ValueError("file too short")
        a = np.frombuffer(raw[off:off + nb].tobytes(), dtype)
        off += nb
        return a

    cp = take("<i8", nc + 1)
    rl = take("<u2", n)
    col = take("<u4", stored)
    if hd["packed"]:
        diag = take("<f8", n * vs)
        vt_re = take("<f8", hd["n_values"])
        vt_im = take("<f8", hd["n_values_im"]) if cplx else None
    else:
        vals = take("<f8", stored * vs)
    if off != len(raw):
        raise ValueError("trailing bytes")
    if cp[0] != 0 or cp[-1] != stored or np.any(np.diff(cp) % 64) or np.any(np.diff(cp) < 0):
        raise ValueError("bad chunk offsets")
    ind = np.zeros(n * W, np.uint32)
    val = np.zeros(n * W, np.complex128 if cplx else np.float64)
    track = rl.astype(np.int16)
    rows = np.arange(n, dtype=np.int64)
    widths = (np.diff(cp) // 64).astype(np.int64)
    lane = rows % 64
    base = cp[rows // 64] + lane
    if np.any(rl.astype(np.int64) > widths[rows // 64]) or np.any(rl.astype(np.int64) > W):
        raise ValueError("row longer than its chunk")
    ind[rows * W] = rows.astype(np.uint32)  # slot 0 column of every row (empty rows too)
    cb, rb = hd["col_bits"], hd["re_bits"]
    for s in range(int(rl.max()) if n else 0):
        live = rl > s
        r = rows[live]
        pos = base[live] + 64 * s
        w = col[pos]
        if hd["packed"]:
            if s == 0:
                c = r.astype(np.uint32)
                v = _cplx(diag[2 * r], diag[2 * r + 1]) if cplx else diag[r]
            else:
                c = w & np.uint32((1 << cb) - 1)
                if cplx:
                    v = _cplx(vt_re[(w >> np.uint32(cb)) & np.uint32((1 << rb) - 1)], vt_im[w >> np.uint32(cb + rb)])
                else:
                    v = vt_re[w >> np.uint32(cb)]
        else:
            c = w
            v = _cplx(vals[2 * pos], vals[2 * pos + 1]) if cplx else vals[pos]
        if np.any(c >= n):
            raise ValueError("column out of range")
        ind[r * W + s] = c
        val[r * W + s] = v
    return hd, ind, val, track
