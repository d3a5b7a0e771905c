"""MPS reader on the CPU: the restatement against the reference's own tests
(cpp/mps_reader/tests/test_mpsreader.cpp, transcribed in
tests/golden/reference_mps.json) and examples/model.mps; against the
reference reader itself on seeded random texts where oracle/_ref holds its
build; and the engine's host parser (gp_mps_parse, no GPU call) against the
restatement on the same texts, errors and messages included."""
import json
import math
import os

import numpy as np
import pytest

import gpupre as G
import mps_texts as T
import oracle_mps as M

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def kat_cases():
    with open(os.path.join(GOLDEN, "reference_mps.json")) as f:
        cases = json.load(f)["cases"]
    for case in cases:
        if "file" in case:
            with open(os.path.join(GOLDEN, case["file"]), "rb") as f:
                case["text"] = f.read()
        else:
            case["text"] = case["text"].encode()
    return cases


def _f(v):
    return float(v) if not isinstance(v, str) else float(v.replace("inf", "Infinity"))


def check_expect(lp, exp):
    if "u_len" in exp:
        assert len(lp.u) == exp["u_len"] and lp.u[0] > exp["u0_greater_than"]
        return
    assert lp.name == exp["name"].encode()
    assert (lp.m, lp.n) == (exp["m"], exp["n"])
    for k in ("c", "AL", "AU", "l", "u", "csc_val"):
        if k in exp:
            assert np.array_equal(np.asarray(lp.__dict__[k]), np.array([_f(v) for v in exp[k]])), k
    for k in ("col_ptr", "row_idx"):
        if k in exp:
            assert list(lp.__dict__[k]) == exp[k], k
    if "obj_constant" in exp:
        assert lp.obj_constant == exp["obj_constant"]


@pytest.mark.parametrize("case", kat_cases(), ids=lambda c: c["test"])
def test_restatement_reference_known_answers(case):
    check_expect(M.read_mps(case["text"]), case["expect"])


def test_micro_bound_is_the_running_product():
    # 0.000001 read digit by digit: 1 * (0.1 * 0.1 * ... six factors), a
    # double above 1e-6 -- the rounding the reference pins
    lp = M.read_mps(kat_cases()[2]["text"])
    p = 0.1
    for _ in range(5):
        p *= 0.1
    assert lp.u[0] == p and p > 1e-6 and p != 1e-6


@pytest.mark.parametrize("token,value", [("7", 7.0), ("-12", -12.0), ("+3", 3.0), ("2.5", 2.5), ("1e3", 1000.0),
                                         ("1E+3", 1000.0), ("2d2", 200.0), ("-4D-1", -0.4), ("inf", math.inf),
                                         ("-Infinity", -math.inf), ("0x1p3", 8.0), ("5.", 5.0)])
def test_number_spellings_have_known_values(token, value):
    # digit path: integers and x.5 are exact; exponents scale by pow(10, e);
    # 'd' is an exponent letter; the rest goes to strtod
    assert M.number(token.encode()) == value


def test_number_long_mantissa_uses_the_library_reading():
    tok = b"0.1234567890123456789"  # 19 fraction digits
    assert M.number(tok) == float("0.1234567890123456789")


@pytest.mark.parametrize("text,message", [
    (b"NAME x\nOBJSENSE\n    MAX\nROWS\n N c\nENDATA\n", "MPS OBJSENSE MAX is not supported by the MPS reader"),
    (b"NAME x\nROWS\n N c\nQUADOBJ\nENDATA\n", "quadratic MPS sections are not supported by the MPS reader"),
    (b"NAME x\nROWS\n N c\n E r\nCOLUMNS\n    x r 1 q 2\nENDATA\n", "unknown MPS row in COLUMNS: q"),
    (b"NAME x\nROWS\n N c\n E r\nCOLUMNS\n    x r 1\nBOUNDS\n UP b x\nENDATA\n",
     "missing value for MPS bound type: UP"),
    (b"NAME x\nROWS\n N c\n Z r\nENDATA\n", "unsupported MPS row type: Z"),
    (b"NAME x\nROWS\n N c\nENDATA\n", "empty MPS model"),
    (b"    x r 1\n", "MPS data entry outside a supported section"),
    (b"NAME x\nFOO\n", "unknown MPS section header: FOO"),
    (b"NAME x\nROWS\n N c\n E r\nCOLUMNS\n    x r 1.2.3\nENDATA\n", "invalid numeric token in MPS: 1.2.3"),
    (b"NAME x\nROWS\n N c\n E r\nCOLUMNS\n    x r abc\nENDATA\n", "stod"),
    (b"NAME x\nROWS\n N\nENDATA\n", "invalid ROWS entry at line 3"),
])
def test_errors_carry_the_reference_messages(text, message):
    with pytest.raises(M.MpsReadError) as e:
        M.read_mps(text)
    assert str(e.value) == message
    with pytest.raises(G.GpError) as g:
        G.parse_mps(text)
    assert g.value.code == 1 and str(g.value).split(": ", 1)[1] == message


@pytest.mark.skipif(not M.reference_available(), reason="oracle/_ref/libmpsref.so not built (no reference tree)")
@pytest.mark.parametrize("block", range(4))
def test_restatement_matches_reference_reader(block):
    for seed in range(block * 100, block * 100 + 100):
        text = T.random_text(seed)
        try:
            ours, our_err = M.read_mps(text), None
        except M.MpsReadError as e:
            ours, our_err = None, str(e)
        try:
            ref, ref_err = M.reference_read_mps(text), None
        except M.MpsReadError as e:
            ref, ref_err = None, str(e)
        assert our_err == ref_err, seed
        if ours is not None:
            assert M.same(ours, ref) == [], seed


def host_csc(model: G.MpsModel):
    """The parser's file-order entries put in the reference's CSC by a plain
    loop: stable by (column, row), duplicates summed left to right from 0."""
    a = model.arrays()
    r, c, v = a["entry_row"], a["entry_col"], a["entry_val"]
    order = np.lexsort((r, c))
    col_ptr, row_idx, vals = [0], [], []
    k = 0
    for j in range(model.n):
        while k < len(order) and c[order[k]] == j:
            i, s = r[order[k]], 0.0
            while k < len(order) and c[order[k]] == j and r[order[k]] == i:
                s += float(v[order[k]])
                k += 1
            row_idx.append(i)
            vals.append(s)
        col_ptr.append(len(vals))
    return np.array(col_ptr, np.int64), np.array(row_idx, np.int32), np.array(vals, np.float64), a


def same_bits(x, y):
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    return x.shape == y.shape and bool(np.all((x.view(np.uint64) == y.view(np.uint64)) | (np.isnan(x) & np.isnan(y))))


@pytest.mark.parametrize("block", range(4))
def test_host_parser_matches_restatement(block):
    for seed in range(block * 100, block * 100 + 100):
        text = T.random_text(seed)
        try:
            ref, ref_err = M.read_mps(text), None
        except M.MpsReadError as e:
            ref, ref_err = None, str(e)
        try:
            got, got_err = G.parse_mps(text), None
        except G.GpError as e:
            got, got_err = None, str(e).split(": ", 1)[1]
            assert e.code == 1
        assert got_err == ref_err, seed
        if got is None:
            continue
        try:
            col_ptr, row_idx, vals, a = host_csc(got)
            assert (got.name, got.m, got.n) == (ref.name, ref.m, ref.n), seed
            assert np.array_equal(col_ptr, ref.col_ptr) and np.array_equal(row_idx, ref.row_idx), seed
            assert same_bits(vals, ref.csc_val), seed
            for k in ("c", "AL", "AU", "l", "u"):
                assert same_bits(a[k], getattr(ref, k)), (seed, k)
            assert same_bits([got.obj_constant], [ref.obj_constant]), seed
        finally:
            got.free()


def test_read_file_and_missing_file(tmp_path):
    with G.read_mps(os.path.join(GOLDEN, "model.mps")) as m:
        assert (m.m, m.n, m.entries) == (2, 2, 4)
        assert list(m.arrays()["c"]) == [-3.0, -5.0]
    with pytest.raises(G.GpError) as e:
        G.read_mps(str(tmp_path / "absent.mps"))
    assert e.value.code == 3 and "failed to open" in str(e.value)


def test_integer_writer_reads_back_exactly():
    rng = np.random.default_rng(3)
    text, exp = T.write_integer_lp(rng, 300, 400, 2500)
    ref = M.read_mps(text)
    order = np.lexsort((ref.col_idx, np.repeat(np.arange(ref.m), np.diff(ref.row_ptr))))
    assert np.array_equal(ref.row_ptr, exp["row_ptr"]) and np.array_equal(ref.col_idx, exp["col_idx"])
    assert np.array_equal(ref.csr_val[order], exp["vals"])
    for k in ("c", "AL", "AU", "l", "u"):
        assert np.array_equal(getattr(ref, k), exp[k]), k


def test_package_writer_reads_back():
    # gpupre.write_mps: structure and infinities exactly; FP64 values to the
    # digit path's rounding (a few ulp); integer-valued matrix and sides exactly
    lp = G.generate_workload("small", seed=0)
    r = M.read_mps(G.write_mps(lp))
    assert (r.m, r.n) == (lp.m, lp.n)
    assert np.array_equal(r.row_ptr, lp.row_ptr) and np.array_equal(r.col_idx, lp.col_idx)
    assert np.allclose(r.csr_val, lp.vals, rtol=1e-14, atol=0)
    for k in ("AL", "AU", "l", "u", "c"):
        a, b = getattr(r, k), np.asarray(getattr(lp, k))
        assert np.array_equal(np.isfinite(a), np.isfinite(b)), k
        f = np.isfinite(b)
        assert np.allclose(a[f], b[f], rtol=1e-14, atol=0), k
    ilp = G.generate(1500, 2000, 3, degree_mode=0, mean_degree=5.0, integral=1)
    r = M.read_mps(G.write_mps(ilp))
    assert np.array_equal(r.csr_val, ilp.vals)
    for k in ("AL", "AU", "l", "u"):  # the recipe's costs are not integers
        assert np.array_equal(getattr(r, k), np.asarray(getattr(ilp, k))), k
