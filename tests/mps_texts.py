"""MPS texts for the ingest tests: seeded random texts that exercise every
card form of the reference reader (free and fixed format, comments, CRLF,
markers, ranges, all bound types, number spellings that take the digit
path and the C-library path, and one planted defect in some of them), and a
writer of large integer LPs whose reading is exact.
"""
from __future__ import annotations

import numpy as np

_NUMBER_FORMS = ("int", "fixed3", "g6", "repr", "exp", "expd", "plus", "lead_dot", "trail_dot", "long",
                 "special", "hex")


def _number(rng, v: float) -> str:
    form = _NUMBER_FORMS[rng.integers(len(_NUMBER_FORMS))]
    if form == "int":
        return str(int(round(v)))
    if form == "fixed3":
        return f"{v:.3f}"
    if form == "g6":
        return f"{v:.6g}"
    if form == "repr":
        return repr(float(v))
    if form == "exp":
        return f"{v:.4e}".replace("e", "eE"[rng.integers(2)])
    if form == "expd":
        return f"{v:.3e}".replace("e", "dD"[rng.integers(2)])
    if form == "plus":
        return f"+{abs(v):.2f}"
    if form == "lead_dot":
        s = f"{abs(v) % 1:.4f}"[1:]
        return ("-" if v < 0 else "") + s
    if form == "trail_dot":
        return f"{int(v)}."
    if form == "long":  # more than 18 mantissa digits: the C-library path
        return f"{v:.15f}1234"
    if form == "special":
        return ("inf", "-inf", "Infinity", "1e400", "-1e-400", "0", "-0", "0.000001", "1e+2", "1E-3")[
            rng.integers(10)]
    return float(v).hex()


def _name(prefix: str, k: int, fixed: bool, spaced: bool) -> str:
    if fixed and spaced:
        return f"{prefix} {k}"[:8]
    return f"{prefix}{k}"


def _fixed_line(f1: str, f2: str = "", f3: str = "", f4: str = "", f5: str = "", f6: str = "") -> str:
    """Data card in fixed columns 2-3, 5-12, 15-22, 25-36, 40-47, 50-61."""
    line = " " + f1.ljust(2)[:2] + " " + f2.ljust(8)[:8] + "  " + f3.ljust(8)[:8] + "  " + f4.ljust(12)[:12]
    if f5 or f6:
        line += "   " + f5.ljust(8)[:8] + "  " + f6.ljust(12)[:12]
    return line.rstrip()


def random_text(seed: int, rows: int = 12, cols: int = 15, defect_rate: float = 0.3) -> bytes:
    """A random MPS text; in about `defect_rate` of the seeds one planted defect."""
    rng = np.random.default_rng(seed)
    fixed = bool(rng.integers(3) == 0)
    spaced = fixed and bool(rng.integers(2))  # names with blanks: the free reading fails
    eol = "\r\n" if rng.integers(4) == 0 else "\n"
    out = []

    def card(*fields):
        """fields[0] is the type field (columns 2-3 in fixed format), "" when none."""
        if fixed:
            out.append(_fixed_line(*fields))
        else:
            sep = lambda: " " * int(rng.integers(1, 4)) if rng.integers(5) else "\t"  # noqa: E731
            line = " " * int(rng.integers(1, 5))
            for f in fields:
                if f:
                    line += f + sep()
            out.append(line.rstrip())

    def noise():
        r = rng.integers(12)
        if r == 0:
            out.append("* a comment line")
        elif r == 1:
            out.append("")
        elif r == 2:
            out.append("   & continuation-style comment")

    defect = rng.random() < defect_rate
    kind = int(rng.integers(14)) if defect else -1
    if kind == 0:
        card("", "X1", "R1", "1")  # data before any section
    out.append("NAME" + (("          " if fixed else " ") + f"model{seed}" if rng.integers(4) else ""))
    if rng.integers(3) == 0:
        out.append("OBJSENSE")
        card("", "MAX" if kind == 1 else ("MIN" if rng.integers(2) else "MAXIMIZE"))
    out.append("ROWS")
    rnames = [_name("R", i, fixed, spaced) for i in range(rows)]
    types = [("E", "G", "L")[rng.integers(3)] for _ in range(rows)]
    card("N", "COST")
    extra_n = bool(rng.integers(2))
    for i in range(rows):
        noise()
        t = types[i]
        if kind == 2 and i == rows // 2:
            t = "X"
        card(t if rng.integers(4) else {"E": "EQ", "G": "GE", "L": "LE", "X": "X"}[t], rnames[i])
        if extra_n and i == 1:
            card("N", "FREE1")
    if rng.integers(4) == 0 and not fixed:
        card("E", rnames[0])  # a repeated row name keeps its first meaning
    if kind == 3:
        out.append("QUADOBJ")
    out.append("COLUMNS")
    cnames = [_name("C", j, fixed, spaced) for j in range(cols)]
    marker = False
    for j in range(cols):
        if rng.integers(6) == 0:
            marker = not marker
            card("", f"M{j}", "'MARKER'", "'INTORG'" if marker else "'INTEND'")
        rowset = rng.choice(rows, size=int(rng.integers(0, min(rows, 5) + 1)), replace=False)
        pairs = [("COST", _number(rng, rng.normal() * 3))] if rng.integers(3) else []
        for i in rowset:
            pairs.append((rnames[i], _number(rng, rng.normal() * 4)))
            if rng.integers(5) == 0:  # a duplicate entry
                pairs.append((rnames[i], _number(rng, rng.normal())))
        if extra_n and rng.integers(3) == 0:
            pairs.append(("FREE1", "7"))
        if kind == 4 and j == cols // 2:
            pairs.append(("NOSUCHROW", "1"))
        if kind == 5 and j == 0:
            pairs.append((rnames[0], "1.2.3"))
        if not pairs:
            pairs = [("COST", "0")]
        for k in range(0, len(pairs), 2):
            noise()
            p = pairs[k:k + 2]
            if len(p) == 2 and rng.integers(2):
                card("", cnames[j], p[0][0], p[0][1], p[1][0], p[1][1])
            else:
                card("", cnames[j], p[0][0], p[0][1])
                if len(p) == 2:
                    card("", cnames[j], p[1][0], p[1][1])
    if marker:
        card("", "MEND", "'MARKER'", "'INTEND'")
    if kind == 6:
        out.append("SOMETHING")
    out.append("RHS")
    for vec in ("RHS1", "RHS2")[:int(rng.integers(1, 3))]:
        for i in rng.choice(rows, size=rows // 2, replace=False):
            card("", vec, rnames[i], _number(rng, rng.normal() * 10))
        if rng.integers(2):
            card("", vec, "COST", _number(rng, rng.normal()))
    if kind == 7:
        card("", "RHS1")
    if rng.integers(2):
        out.append("RANGES")
        for vec in ("RNG1", "RNG2")[:int(rng.integers(1, 3))]:
            for i in rng.choice(rows, size=rows // 3, replace=False):
                card("", vec, rnames[i], _number(rng, rng.normal() * 5))
    out.append("BOUNDS")
    btypes = ("LO", "UP", "FX", "FR", "MI", "PL", "BV", "LI", "UI")
    for vec in ("BND1", "BND2")[:int(rng.integers(1, 3))]:
        for j in rng.choice(cols, size=cols // 2, replace=False):
            t = btypes[rng.integers(len(btypes))]
            if t in ("FR", "MI", "PL", "BV"):
                card(t, vec, cnames[j])
            else:
                v = rng.normal() * 5
                if t == "UP" and rng.integers(3) == 0:
                    v = -abs(v)  # a lone negative upper bound
                card(t, vec, cnames[j], _number(rng, v))
    if kind == 8:
        card("UP", "BND1", cnames[0])
    if kind == 9:
        card("SC", "BND1", cnames[0], "4")
    if kind == 10:
        card("LO", "BND1", "NOCOL", "1")
    if kind == 11:
        out.append("OBJECT BOUND")
    if kind == 12:
        card("R")
    out.append("ENDATA")
    if kind == 13:
        out = ["NAME empty", "ROWS", " N COST", "ENDATA"]
    if rng.integers(3) == 0:
        out.append("    junk after the end")
    text = eol.join(out)
    if rng.integers(4):
        text += eol
    return text.encode("latin-1")


def write_integer_lp(rng, m: int, n: int, nnz: int, dup_rate: float = 0.05):
    """A random LP with small integer data written as a free-format MPS text.

    Every matrix value is a nonzero integer; a share of the entries is split
    into two cards that sum to it exactly.  Returns (text, expected) where
    expected holds the host CSR and vectors the reading must give."""
    flat = np.unique(rng.integers(0, m * n, size=nnz, dtype=np.int64))
    rows = (flat // n).astype(np.int32)
    cols = (flat % n).astype(np.int32)
    vals = rng.integers(1, 10, size=flat.size) * rng.choice([-1, 1], size=flat.size)
    c = rng.integers(-5, 6, size=n).astype(np.float64)
    kind = rng.integers(0, 4, size=m)  # 0 E, 1 G, 2 L, 3 ranged L
    rhs = rng.integers(-20, 21, size=m).astype(np.float64)
    rng_w = rng.integers(1, 9, size=m).astype(np.float64)
    AL = np.where(kind == 0, rhs, np.where(kind == 1, rhs, np.where(kind == 3, rhs - rng_w, -np.inf)))
    AU = np.where(kind == 0, rhs, np.where(kind == 1, np.inf, rhs))
    btype = rng.integers(0, 5, size=n)  # 0 default, 1 LO+UP, 2 FR, 3 MI+UP, 4 FX
    lo = rng.integers(-5, 1, size=n).astype(np.float64)
    hi = lo + rng.integers(0, 6, size=n)
    l = np.where(btype == 0, 0.0, np.where(btype == 1, lo, np.where(btype == 4, lo, -np.inf)))
    u = np.where(btype == 0, np.inf, np.where(btype == 1, hi, np.where(btype == 2, np.inf,
                                                                      np.where(btype == 4, lo, hi))))
    lines = ["NAME intlp", "ROWS", " N obj"]
    tchar = np.array(["E", "G", "L", "L"])[kind]
    lines += [f" {t} r{i}" for i, t in enumerate(tchar)]
    lines.append("COLUMNS")
    order = np.lexsort((rows, cols))  # column-major, as MPS files list entries
    ptr = np.searchsorted(cols[order], np.arange(n + 1))
    split = rng.random(flat.size) < dup_rate
    for j in range(n):
        if c[j] != 0 or ptr[j] == ptr[j + 1]:  # every column needs a card to exist
            lines.append(f"    x{j} obj {int(c[j])}")
        for q in order[ptr[j]:ptr[j + 1]]:
            v = int(vals[q])
            if split[q]:
                a = int(rng.integers(-9, 10))
                lines.append(f"    x{j} r{rows[q]} {a} r{rows[q]} {v - a}")
            else:
                lines.append(f"    x{j} r{rows[q]} {v}")
    lines.append("RHS")
    lines += [f"    rhs r{i} {int(rhs[i])}" for i in range(m) if rhs[i] != 0]
    lines.append("RANGES")
    lines += [f"    rng r{i} {int(rng_w[i])}" for i in range(m) if kind[i] == 3]
    lines.append("BOUNDS")
    for j in range(n):
        b = btype[j]
        if b == 1:
            lines += [f" LO bnd x{j} {int(lo[j])}", f" UP bnd x{j} {int(hi[j])}"]
        elif b == 2:
            lines.append(f" FR bnd x{j}")
        elif b == 3:
            lines += [f" MI bnd x{j}", f" UP bnd x{j} {int(hi[j])}"]
        elif b == 4:
            lines.append(f" FX bnd x{j} {int(lo[j])}")
    lines.append("ENDATA")
    csr_order = np.lexsort((cols, rows))
    row_ptr = np.searchsorted(rows[csr_order], np.arange(m + 1)).astype(np.int32)
    expected = {"m": m, "n": n, "row_ptr": row_ptr, "col_idx": cols[csr_order], "vals": vals[csr_order].astype(
        np.float64), "c": c, "AL": AL, "AU": AU, "l": l, "u": u}
    return ("\n".join(lines) + "\n").encode(), expected
