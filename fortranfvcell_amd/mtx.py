"""MatrixMarket interchange of solver systems, in the layout the reference
exports them (src/ExportCoo.f:23-62 with src/CsrToCoo.f:29-179).

The reference writes one coordinate file per system, `<name>.mtx`:

    %%MatrixMarket matrix coordinate real general      (symmetric when not unsym)
    neq neq nnz                                         (3i9)
    lin col val                                         (2i9,e24.16), row by row:
                                                         the diagonal first, then
                                                         the row's off-diagonals
                                                         in storage order

and the right-hand side as an array file `<name>_b.mtx` ((2i9) header, then
one (e24.16) value per line).  `write_system` produces exactly that from any of
the layouts in problems.to_ell / to_csrd / to_csrc (ELLPACK padding slots are
skipped); `exact=True` writes 17 significant digits instead of E24.16's 16, so
that values round-trip bit for bit.  The values are kept in FP64 throughout
(the reference allocates its COO value buffer with 4-byte elements,
src/ExportCoo.f:44, which this does not reproduce).

`read_system` reads such a pair (or any coordinate real general/symmetric
file) back into the CSRD layout of problems.to_csrd, ready for the solver.
"""
from __future__ import annotations

import os

import numpy as np


def fortran_e(x: float, w: int = 24, d: int = 16) -> str:
    """Fortran's Ew.d edit descriptor (mantissa in [0.1, 1), two-digit
    exponent while it fits), right-justified in w columns."""
    if x == 0.0 or not np.isfinite(x):
        body = f"0.{'0' * d}E+00" if x == 0.0 else str(x)
        return body.rjust(w)
    m, e = f"{abs(x):.{d - 1}e}".split("e")
    digits = m.replace(".", "")
    exp = int(e) + 1
    mant = f"0.{digits}"
    sign = "+" if exp >= 0 else "-"
    # two-digit exponents keep the E; |exp| > 99 drops it (Fortran Ew.d)
    tail = f"E{sign}{abs(exp):02d}" if abs(exp) <= 99 else f"{sign}{abs(exp):03d}"
    body = f"{'-' if x < 0 else ''}{mant}{tail}"
    return body.rjust(w)


def _rows(fmt: dict):
    """(row, col, value) triples, 1-based, in the reference's export order."""
    neq = int(fmt["neq"])
    m = int(fmt["matriz"])
    ad = np.asarray(fmt["ad"], np.float64)
    ja = np.asarray(fmt["ja"], np.int64)
    al = np.asarray(fmt["al"], np.float64)
    au = al if fmt.get("au") is None else np.asarray(fmt["au"], np.float64)
    lin, col, val = [], [], []
    if m == 4:                                   # ELLPACK: ja(w, neq) slot-fastest
        w = int(fmt["iax"][0])
        a2, j2 = al.reshape(neq, w), ja.reshape(neq, w)
        for i in range(neq):
            lin.append(i + 1), col.append(i + 1), val.append(ad[i])
            for s in range(w):
                if a2[i, s] != 0.0 or j2[i, s] != 1:
                    lin.append(i + 1), col.append(int(j2[i, s])), val.append(a2[i, s])
    elif m == 1:                                 # CSRD
        ia = np.asarray(fmt["iax"], np.int64)
        for i in range(neq):
            lin.append(i + 1), col.append(i + 1), val.append(ad[i])
            for k in range(ia[i] - 1, ia[i + 1] - 1):
                lin.append(i + 1), col.append(int(ja[k])), val.append(al[k])
    elif m == 3:                                 # CSRC: lower pattern, au = transposed upper
        # The reference's export walks only the stored lower pattern
        # (src/CsrToCoo.f:84-107; its col > nl branch never fires there), so a
        # nonsymmetric CSRC matrix would lose its upper triangle; here row j
        # also gets its upper entries (j, i) = au(k) from the rows i > j
        # whose lists hold column j, in ascending column order.
        ia = np.asarray(fmt["iax"], np.int64)
        upper = [[] for _ in range(neq)]
        unsym = not fmt.get("sym", False) and fmt.get("au") is not None
        if unsym:
            for i in range(neq):
                for k in range(ia[i] - 1, ia[i + 1] - 1):
                    upper[int(ja[k]) - 1].append((i + 1, au[k]))
        for i in range(neq):
            lin.append(i + 1), col.append(i + 1), val.append(ad[i])
            for k in range(ia[i] - 1, ia[i + 1] - 1):
                lin.append(i + 1), col.append(int(ja[k])), val.append(al[k])
            for c, v in sorted(upper[i]):
                lin.append(i + 1), col.append(c), val.append(v)
    elif m == 2:                                 # full CSR: the row as stored
        ia = np.asarray(fmt["iax"], np.int64)
        for i in range(neq):
            for k in range(ia[i] - 1, ia[i + 1] - 1):
                lin.append(i + 1), col.append(int(ja[k])), val.append(ad[k])
    else:
        raise ValueError(f"matriz {m}")
    return np.array(lin), np.array(col), np.array(val)


def write_system(fmt: dict, path_prefix: str, unsym: bool | None = None, exact: bool = False) -> tuple:
    """Write `<prefix>.mtx` and `<prefix>_b.mtx`; returns the two paths."""
    unsym = (not fmt.get("sym", False)) if unsym is None else unsym
    lin, col, val = _rows(fmt)
    neq = int(fmt["neq"])
    num = (lambda v: f"{v: .16e}".rjust(24)) if exact else fortran_e
    head = "%%MatrixMarket matrix coordinate real " + ("general" if unsym else "symmetric")
    a_path, b_path = path_prefix + ".mtx", path_prefix + "_b.mtx"
    with open(a_path, "w") as fh:
        fh.write(head + "\n")
        fh.write(f"{neq:9d}{neq:9d}{len(val):9d}\n")
        for r, c, v in zip(lin, col, val):
            fh.write(f"{r:9d}{c:9d}{num(v)}\n")
    with open(b_path, "w") as fh:
        fh.write("%%MatrixMarket matrix array real general\n")
        fh.write(f"{neq:9d}{1:9d}\n")
        for v in np.asarray(fmt["b"], np.float64):
            fh.write(num(v) + "\n")
    return a_path, b_path


def _float(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def read_system(path_prefix: str) -> dict:
    """Read `<prefix>.mtx` (+ `<prefix>_b.mtx` when present) into the CSRD
    layout (matriz 1): ad, ia(neq+1), ja(nad) without the diagonal and sorted
    within each row, al = au holding both triangles (the symmetric header
    mirrors the lower entries)."""
    with open(path_prefix + ".mtx") as fh:
        head = fh.readline().lower()
        if "coordinate" not in head or "real" not in head:
            raise ValueError(f"not a real coordinate MatrixMarket file: {head.strip()}")
        symmetric = "symmetric" in head
        line = fh.readline()
        while line.startswith("%"):
            line = fh.readline()
        n, n2, nnz = (int(t) for t in line.split())
        if n != n2:
            raise ValueError("the matrix is not square")
        data = np.loadtxt(fh, dtype=str, ndmin=2) if nnz else np.zeros((0, 3), dtype=str)
    r = data[:, 0].astype(np.int64)
    c = data[:, 1].astype(np.int64)
    v = np.array([_float(t) for t in data[:, 2]])
    if symmetric:
        off = r != c
        r, c, v = np.concatenate([r, c[off]]), np.concatenate([c, r[off]]), np.concatenate([v, v[off]])
    ad = np.zeros(n)
    diag = r == c
    np.add.at(ad, r[diag] - 1, v[diag])
    ro, co, vo = r[~diag], c[~diag], v[~diag]
    order = np.lexsort((co, ro))
    ro, co, vo = ro[order], co[order], vo[order]
    ia = np.concatenate([[1], 1 + np.cumsum(np.bincount(ro - 1, minlength=n))]).astype(np.int32)
    b = np.zeros(n)
    if os.path.exists(path_prefix + "_b.mtx"):
        with open(path_prefix + "_b.mtx") as fh:
            fh.readline()
            line = fh.readline()
            while line.startswith("%"):
                line = fh.readline()
            b = np.array([_float(t) for t in fh.read().split()])
    return dict(matriz=1, neq=n, nad=int(len(vo)), ad=ad, iax=ia, ja=co.astype(np.int32),
                al=vo, au=None, b=b, sym=False)
