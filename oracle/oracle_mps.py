"""Restatement of the reference's MPS reader in Python (pure loops: small texts).

TEST INFRASTRUCTURE ONLY: the parity checker of the engine's MPS ingest
(gp_mps_parse / gp_mps_upload).  Only the tests may import it.

Follows mps_reader::read_mps (cpp/mps_reader/src/io/mpsreader.cpp) and the
CSC -> CSR conversion (cpp/mps_reader/src/core/sparse_matrix.cpp:111-135);
each function cites the lines it restates.  Text is handled as bytes and
"blank" is the C locale's isspace set, as in the reference.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(HERE, "_ref", "libmpsref.so")

_BLANK = b" \t\n\v\f\r"
_INF = math.inf


class MpsReadError(ValueError):
    """read_mps's std::invalid_argument (the free-format message, mpsreader.cpp:700-707)."""


@dataclass
class MpsLp:
    name: bytes
    m: int
    n: int
    col_ptr: np.ndarray  # CSC (A^T as CSR): n + 1
    row_idx: np.ndarray
    csc_val: np.ndarray
    row_ptr: np.ndarray  # CSR: m + 1
    col_idx: np.ndarray
    csr_val: np.ndarray
    c: np.ndarray
    AL: np.ndarray
    AU: np.ndarray
    l: np.ndarray
    u: np.ndarray
    obj_constant: float


# ---- cards (mpsreader.cpp:73-188) ------------------------------------------------
def _strip(b: bytes) -> bytes:
    return b.strip(_BLANK)


def _words(line: bytes) -> list:
    out, p, n = [], 0, len(line)
    while p < n:
        while p < n and line[p] in _BLANK:
            p += 1
        q = p
        while q < n and line[q] not in _BLANK:
            q += 1
        if q > p:
            out.append(line[p:q])
        p = q
    return out


def _slice(line: bytes, begin: int, end: int) -> bytes:
    """field_slice (mpsreader.cpp:87-94)."""
    if len(line) <= begin:
        return b""
    return _strip(line[begin:min(len(line), end)])


def _comment(line: bytes) -> bool:
    """is_blank_or_comment (mpsreader.cpp:117-125)."""
    for ch in line:
        if ch in _BLANK:
            continue
        return ch in b"*&"
    return True


def _card(line: bytes, fixed: bool):
    """read_card_free / read_card_fixed (mpsreader.cpp:134-188): (header, fields) or None."""
    if _comment(line):
        return None
    header = line[0] not in _BLANK
    if header or not fixed:
        fields = _words(line)
        if header:
            if len(fields) >= 2 and fields[0] == b"OBJECT" and fields[1] == b"BOUND":
                fields = [b"OBJECT BOUND"]
            elif fixed and fields and fields[0] == b"NAME":
                nm = _slice(line, 14, len(line))
                if nm:
                    fields = [b"NAME", nm]
        return header, fields
    fields = [f for f in (_slice(line, 1, 3), _slice(line, 4, 12), _slice(line, 14, 22),
                          _slice(line, 24, 36), _slice(line, 39, 47), _slice(line, 49, 61)) if f]
    return False, fields


# ---- numbers (mpsreader.cpp:191-298) ------------------------------------------------
_libc = C.CDLL(None, use_errno=True)
_libc.strtod.restype = C.c_double
_libc.strtod.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]
_ERANGE = 34


def _stod(token: bytes) -> float:
    """parse_double_fallback (mpsreader.cpp:191-203) with std::stod's failures."""
    s = token.replace(b"d", b"E").replace(b"D", b"E")
    buf = C.create_string_buffer(s)
    end = C.c_char_p()
    C.set_errno(0)
    v = _libc.strtod(buf, C.byref(end))
    consumed = C.cast(end, C.c_void_p).value - C.addressof(buf)
    if consumed == 0:
        raise _Fail("stod")
    if C.get_errno() == _ERANGE:
        raise _Fail("stod")
    if consumed != len(s):
        raise _Fail("invalid numeric token in MPS: " + token.decode("latin-1"))
    return v


def number(token: bytes) -> float:
    """parse_double (mpsreader.cpp:205-298)."""
    n, i = len(token), 0
    negative = False
    if i < n and token[i:i + 1] in (b"-", b"+"):
        negative = token[i:i + 1] == b"-"
        i += 1
    value, scale, digits = 0.0, 0.1, 0
    while i < n and 48 <= token[i] <= 57:
        value = value * 10.0 + float(token[i] - 48)
        digits += 1
        if digits > 18:
            return _stod(token)
        i += 1
    if i < n and token[i:i + 1] == b".":
        i += 1
        while i < n and 48 <= token[i] <= 57:
            value += float(token[i] - 48) * scale
            scale *= 0.1
            digits += 1
            if digits > 18:
                return _stod(token)
            i += 1
    if digits == 0:
        return _stod(token)
    if i < n:
        if token[i:i + 1] not in (b"e", b"E", b"d", b"D"):
            return _stod(token)
        i += 1
        eneg = False
        if i < n and token[i:i + 1] in (b"-", b"+"):
            eneg = token[i:i + 1] == b"-"
            i += 1
        e, ed = 0, 0
        while i < n:
            if not 48 <= token[i] <= 57:
                return _stod(token)
            e = (10 * e + token[i] - 48) & 0xFFFFFFFF  # the int accumulator wraps
            ed += 1
            i += 1
        if ed == 0:
            return _stod(token)
        e = e - (1 << 32) if e >= (1 << 31) else e
        try:
            p = math.pow(10.0, float(e))  # the C library's pow
        except OverflowError:
            p = _INF  # C pow overflows to inf
        value = value / p if eneg else value * p
    return -value if negative else value


class _Fail(Exception):
    """Any exception of one parse pass (read_mps catches std::exception)."""


# ---- model (mpsreader.cpp:56-70, 300-544) --------------------------------------------
@dataclass
class _Raw:
    name: bytes = b""
    objective: bytes = b""
    rows: dict = field(default_factory=dict)     # name -> (type, index)
    cols: dict = field(default_factory=dict)     # name -> index
    row_types: list = field(default_factory=list)
    entries: list = field(default_factory=list)  # per column: [(row, value)]
    c: list = field(default_factory=list)
    AL: list = field(default_factory=list)
    AU: list = field(default_factory=list)
    l: list = field(default_factory=list)
    u: list = field(default_factory=list)
    marked: list = field(default_factory=list)
    obj_constant: float = 0.0


def _col(raw: _Raw, name: bytes, marked: bool) -> int:
    """ensure_col (mpsreader.cpp:300-315)."""
    if name not in raw.cols:
        raw.cols[name] = len(raw.cols)
        raw.entries.append([])
        raw.c.append(0.0)
        raw.l.append(math.nan)
        raw.u.append(math.nan)
        raw.marked.append(marked)
    return raw.cols[name]


def _row(raw: _Raw, name: bytes, section: str):
    if name not in raw.rows:
        raise _Fail(f"unknown MPS row in {section}: " + name.decode("latin-1"))
    return raw.rows[name]


def _rows_line(raw: _Raw, f: list, line: int):
    """read_rows_line (mpsreader.cpp:345-373)."""
    if len(f) < 2:
        raise _Fail(f"invalid ROWS entry at line {line}")
    t = chr(f[0][0])
    if t == "N":
        if not raw.objective:
            raw.objective = f[1]
        raw.rows.setdefault(f[1], ("N", -1))
        return
    if t not in "EGL":
        raise _Fail("unsupported MPS row type: " + t)
    raw.rows.setdefault(f[1], (t, len(raw.row_types)))
    raw.row_types.append(t)
    raw.AL.append(-_INF if t == "L" else 0.0)
    raw.AU.append(_INF if t == "G" else 0.0)


def _columns_line(raw: _Raw, f: list, line: int, marked: bool):
    """read_columns_line (mpsreader.cpp:375-395)."""
    if len(f) < 3:
        raise _Fail(f"invalid COLUMNS entry at line {line}")
    j = _col(raw, f[0], marked)
    for k in range(1, len(f) - 1, 2):
        v = number(f[k + 1])
        t, i = _row(raw, f[k], "COLUMNS")
        if i < 0:
            if f[k] == raw.objective:
                raw.c[j] += v
            continue
        raw.entries[j].append((i, v))


def _rhs_line(raw: _Raw, f: list, line: int, chosen: list):
    """read_rhs_line (mpsreader.cpp:397-431)."""
    if len(f) < 3:
        raise _Fail(f"invalid RHS entry at line {line}")
    if not chosen[0]:
        chosen[0] = f[0]
    if f[0] != chosen[0]:
        return
    for k in range(1, len(f) - 1, 2):
        t, i = _row(raw, f[k], "RHS")
        if i < 0:
            if f[k] == raw.objective:
                raw.obj_constant = -number(f[k + 1])
            continue
        v = number(f[k + 1])
        if t == "E":
            raw.AL[i] = raw.AU[i] = v
        elif t == "G":
            raw.AL[i] = v
        else:
            raw.AU[i] = v


def _ranges_line(raw: _Raw, f: list, line: int, chosen: list):
    """read_ranges_line / apply_range (mpsreader.cpp:433-474)."""
    if len(f) < 3:
        raise _Fail(f"invalid RANGES entry at line {line}")
    if not chosen[0]:
        chosen[0] = f[0]
    if f[0] != chosen[0]:
        return
    for k in range(1, len(f) - 1, 2):
        v = number(f[k + 1])
        t, i = _row(raw, f[k], "RANGES")
        if i < 0:
            continue
        if t == "E":
            if v >= 0.0:
                raw.AU[i] += v
            else:
                raw.AL[i] += v
        elif t == "L":
            raw.AL[i] = raw.AU[i] - abs(v)
        elif t == "G":
            raw.AU[i] = raw.AL[i] + abs(v)


def _bounds_line(raw: _Raw, f: list, line: int, chosen: list):
    """read_bounds_line (mpsreader.cpp:476-531)."""
    if len(f) < 3:
        raise _Fail(f"invalid BOUNDS entry at line {line}")
    if not chosen[0]:
        chosen[0] = f[1]
    if f[1] != chosen[0]:
        return
    if f[2] not in raw.cols:
        raise _Fail("unknown MPS column in BOUNDS: " + f[2].decode("latin-1"))
    j = raw.cols[f[2]]
    t = f[0]
    if t == b"FR":
        raw.l[j], raw.u[j] = -_INF, _INF
        return
    if t == b"MI":
        raw.l[j] = -_INF
        return
    if t == b"PL":
        raw.u[j] = _INF
        return
    if t == b"BV":
        raw.l[j], raw.u[j] = 0.0, 1.0
        return
    if len(f) < 4:
        raise _Fail("missing value for MPS bound type: " + t.decode("latin-1"))
    v = number(f[3])
    if t in (b"LO", b"LI"):
        raw.l[j] = v
    elif t in (b"UP", b"UI"):
        raw.u[j] = v
    elif t == b"FX":
        raw.l[j] = raw.u[j] = v
    else:
        raise _Fail("unsupported MPS bound type: " + t.decode("latin-1"))


def _finalize(raw: _Raw):
    """finalize_variable_bounds (mpsreader.cpp:533-544); a column whose only
    bound card is BV / LI / UI has a bound set, so its integer type never
    reaches the [0, 1] default."""
    for j in range(len(raw.l)):
        lo, up = math.isnan(raw.l[j]), math.isnan(raw.u[j])
        if lo and up:
            raw.l[j], raw.u[j] = 0.0, (1.0 if raw.marked[j] else _INF)
        elif lo:
            raw.l[j] = -_INF if raw.u[j] < 0.0 else 0.0
        elif up:
            raw.u[j] = _INF


_SECTIONS = {b"OBJSENSE": "sense", b"ROWS": "rows", b"COLUMNS": "columns", b"RHS": "rhs",
             b"BOUNDS": "bounds", b"RANGES": "ranges", b"QUADOBJ": "quad", b"QMATRIX": "quad"}


def _parse(lines: list, fixed: bool) -> _Raw:
    """parse_lines (mpsreader.cpp:546-623)."""
    raw = _Raw()
    section, marked = None, False
    rhs, bnd, rng = [b""], [b""], [b""]
    for ln, line in enumerate(lines, start=1):
        card = _card(line, fixed)
        if card is None or not card[1]:
            continue
        header, f = card
        if header:
            h = f[0]
            if h == b"NAME":
                if len(f) >= 2:
                    raw.name = f[1]
                continue
            if h == b"ENDATA":
                break
            if h not in _SECTIONS and h != b"NAME":
                raise _Fail("unknown MPS section header: " + h.decode("latin-1"))
            section = _SECTIONS[h]
            if section == "quad":
                raise _Fail("quadratic MPS sections are not supported by the MPS reader")
            continue
        if section == "sense":
            if f[0] == b"MAX":
                raise _Fail("MPS OBJSENSE MAX is not supported by the MPS reader")
        elif section == "rows":
            _rows_line(raw, f, ln)
        elif section == "columns":
            if len(f) >= 3 and f[1] == b"'MARKER'":
                if f[2] == b"'INTORG'":
                    marked = True
                elif f[2] == b"'INTEND'":
                    marked = False
                continue
            _columns_line(raw, f, ln, marked)
        elif section == "rhs":
            _rhs_line(raw, f, ln, rhs)
        elif section == "bounds":
            _bounds_line(raw, f, ln, bnd)
        elif section == "ranges":
            _ranges_line(raw, f, ln, rng)
        else:
            raise _Fail("MPS data entry outside a supported section")
    if not raw.row_types and not raw.cols:
        raise _Fail("empty MPS model")
    _finalize(raw)
    return raw


def _build(raw: _Raw) -> MpsLp:
    """build_model (mpsreader.cpp:625-667) and CscMatrix::to_csr (sparse_matrix.cpp:111-135)."""
    m, n = len(raw.row_types), len(raw.entries)
    col_ptr, row_idx, vals = [0], [], []
    for ent in raw.entries:
        ent = sorted(ent, key=lambda e: e[0])  # stable
        b = 0
        while b < len(ent):
            r, s, e = ent[b][0], 0.0, b
            while e < len(ent) and ent[e][0] == r:
                s += ent[e][1]
                e += 1
            row_idx.append(r)
            vals.append(s)
            b = e
        col_ptr.append(len(vals))
    counts = [0] * m
    for r in row_idx:
        counts[r] += 1
    row_ptr = [0]
    for r in range(m):
        row_ptr.append(row_ptr[-1] + counts[r])
    nxt = row_ptr[:-1]
    col_idx, csr_val = [0] * len(vals), [0.0] * len(vals)
    for j in range(n):
        for p in range(col_ptr[j], col_ptr[j + 1]):
            d = nxt[row_idx[p]]
            nxt[row_idx[p]] += 1
            col_idx[d], csr_val[d] = j, vals[p]
    f64 = lambda a: np.array(a, dtype=np.float64)  # noqa: E731
    return MpsLp(raw.name, m, n, np.array(col_ptr, np.int64), np.array(row_idx, np.int32), f64(vals),
                 np.array(row_ptr, np.int64), np.array(col_idx, np.int32), f64(csr_val),
                 f64(raw.c), f64(raw.AL), f64(raw.AU), f64(raw.l), f64(raw.u), float(raw.obj_constant))


def split_lines(text: bytes) -> list:
    """std::getline over the text (mpsreader.cpp:687-690)."""
    lines = text.split(b"\n")
    if lines and lines[-1] == b"":
        lines.pop()
    return lines


def read_mps(text) -> MpsLp:
    """read_mps (mpsreader.cpp:686-709): free format, then fixed format."""
    if isinstance(text, str):
        text = text.encode("latin-1")
    lines = split_lines(text)
    try:
        return _build(_parse(lines, False))
    except (_Fail, ValueError, OverflowError) as free_error:
        message = str(free_error)
        try:
            return _build(_parse(lines, True))
        except (_Fail, ValueError, OverflowError):
            raise MpsReadError(message) from None


# ---- the reference's own reader, where ref.mk built it ---------------------------
class _RefLp(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int64),
                ("col_ptr", C.POINTER(C.c_int64)), ("row_idx", C.POINTER(C.c_int32)),
                ("csc_val", C.POINTER(C.c_double)), ("row_ptr", C.POINTER(C.c_int64)),
                ("col_idx", C.POINTER(C.c_int32)), ("csr_val", C.POINTER(C.c_double)),
                ("c", C.POINTER(C.c_double)), ("AL", C.POINTER(C.c_double)), ("AU", C.POINTER(C.c_double)),
                ("l", C.POINTER(C.c_double)), ("u", C.POINTER(C.c_double)), ("obj_constant", C.c_double),
                ("name", C.c_char_p)]


_ref = None


def reference_available() -> bool:
    return os.path.exists(REF_LIB)


def reference_read_mps(text) -> MpsLp:
    """mps_reader::read_mps itself (oracle/_ref/libmpsref.so)."""
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_LIB)
        _ref.mpsref_read.restype = C.c_int
        _ref.mpsref_read.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(_RefLp), C.c_char_p, C.c_size_t]
        _ref.mpsref_free.argtypes = [C.POINTER(_RefLp)]
    if isinstance(text, str):
        text = text.encode("latin-1")
    out, err = _RefLp(), C.create_string_buffer(4096)
    if _ref.mpsref_read(text, len(text), C.byref(out), err, len(err)):
        raise MpsReadError(err.value.decode("latin-1"))
    try:
        arr = lambda p, k, dt: np.ctypeslib.as_array(p, (k,)).astype(dt).copy() if k else np.zeros(0, dt)  # noqa
        m, n, nnz = out.m, out.n, out.nnz
        return MpsLp(out.name, m, n, arr(out.col_ptr, n + 1, np.int64), arr(out.row_idx, nnz, np.int32),
                     arr(out.csc_val, nnz, np.float64), arr(out.row_ptr, m + 1, np.int64),
                     arr(out.col_idx, nnz, np.int32), arr(out.csr_val, nnz, np.float64),
                     arr(out.c, n, np.float64), arr(out.AL, m, np.float64), arr(out.AU, m, np.float64),
                     arr(out.l, n, np.float64), arr(out.u, n, np.float64), out.obj_constant)
    finally:
        _ref.mpsref_free(C.byref(out))


def same(a: MpsLp, b: MpsLp) -> list:
    """Names of the fields where two readings differ (floats compared by bits)."""
    bad = []
    for k in ("name", "m", "n"):
        if getattr(a, k) != getattr(b, k):
            bad.append(k)
    for k in ("col_ptr", "row_idx", "row_ptr", "col_idx"):
        if not np.array_equal(getattr(a, k), getattr(b, k)):
            bad.append(k)
    for k in ("csc_val", "csr_val", "c", "AL", "AU", "l", "u"):
        x, y = np.asarray(getattr(a, k), np.float64), np.asarray(getattr(b, k), np.float64)
        if x.shape != y.shape or not np.array_equal(x.view(np.uint64), y.view(np.uint64)):
            bad.append(k)
    if np.float64(a.obj_constant).view(np.uint64) != np.float64(b.obj_constant).view(np.uint64):
        bad.append("obj_constant")
    return bad
