"""The device line evaluator (warpjq_amd/csrc/jscan.h), compiled for the
host, against the oracle line by line — the same source the gfx950 kernels
compile, pinned without a GPU. Also the device number parser against glibc
strtod (correctly rounded, the contract of Rust's f64::from_str that
json.rs:73-78 relies on)."""
import ctypes
import json
import os
import random
import struct

import pytest

import corpus
import hostsim
import oracle

AGG = ("count", "sum", "min", "max", "avg", "group")
CONF = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "conformance.json"), encoding="utf-8"))


def _is_agg(q):
    return any(q.startswith(a) or ("| " + a) in q for a in AGG)


def _expected_status(st):
    if st["lines_in"] == 0:
        return 4
    if st["malformed"]:
        return 1
    if st["type_errors"]:
        return 2
    return 0


@pytest.mark.parametrize("label,data", ids=lambda v: v if isinstance(v, str) else "data", argvalues=[
    ("15% malformed", corpus.hostile(2, 2500, 0.15)),
    ("crlf", corpus.hostile(3, 600, 0.05, crlf=True)),
    ("edge", corpus.edge_corpus()),
])
def test_per_line_status_and_rows(label, data):
    lines = data.split(b"\n")
    compared = 0
    for q in corpus.QUERIES:
        for csv in (False, True):
            ev = hostsim.Evaluator(q, csv)
            for ln in lines:
                try:
                    out, st = oracle.run(q, ln + b"\n", fmt=oracle.FMT_CSV if csv else oracle.FMT_NDJSON, threads=1)
                except oracle.RunError:
                    break  # --csv on a whole-line query
                o, row, gk = ev.eval(ln)
                if csv and b"\n" in out:
                    out = out.split(b"\n", 1)[1]
                assert o.status == _expected_status(st), (q, csv, ln)
                if o.status == 0 and not _is_agg(q):
                    assert (row if o.selected else b"") == out, (q, csv, ln)
                    compared += 1
    assert compared > len(lines) // 4


def test_aggregate_values_and_group_keys():
    data = corpus.hostile(9, 1500, 0.05)
    for q, key in [("group_by(.host) | sum(.bytes)", "host"), ("group_by(.flag) | count", "flag"),
                   ("group_by(.msg) | count", "msg"), ("group_by(.n) | count", "n")]:
        ev = hostsim.Evaluator(q)
        for ln in data.split(b"\n"):
            o, _, gk = ev.eval(ln)
            if o.status != 0 or not o.selected:
                continue
            out, _ = oracle.run(q, ln + b"\n", threads=1)
            row = json.loads(out)
            v = row[key]
            if o.gkind == 4:  # string: decoded bytes
                assert gk.decode("utf-8", "surrogatepass") == v or gk == v.encode("utf-8", "surrogatepass")
            elif o.gkind == 1:
                assert v is None
            elif o.gkind == 2:
                assert v == (gk == b"true")
            else:
                assert out.split(b":", 1)[1].rsplit(b",", 1)[0] == gk


def test_conformance_verdicts():
    for name, doc in CONF["must_accept"]:
        assert hostsim.validate(doc.encode()) == 0, name
    for name, doc in CONF["must_reject"]:
        assert hostsim.validate(doc.encode()) == 1, name
    for name, doc, acc in CONF["implementation_defined"]:
        assert hostsim.validate(doc.encode()) == (0 if acc else 1), name


def test_depth_limit_declines():
    # 65 brackets hold 64 open containers plus one empty one: still on device
    assert hostsim.validate(b"[" * 65 + b"]" * 65) == 0
    assert hostsim.validate(b"[" * 66 + b"]" * 66) == 3  # FALLBACK past the 64-level stack
    assert hostsim.validate(b"[" * 66) == 3


_libc = ctypes.CDLL("libc.so.6")
_libc.strtod.restype = ctypes.c_double
_libc.strtod.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def _gen(rng):
    k = rng.randrange(9)
    if k == 0:
        return repr(rng.random() * 10.0 ** rng.randrange(-300, 300))
    if k == 1:
        return "%.17g" % rng.uniform(0, 1e6)
    if k == 2:
        return str(rng.randrange(0, 2 ** 64))
    if k == 3:
        return "%de%d" % (rng.randrange(1, 10 ** rng.randrange(1, 25)), rng.randrange(-350, 320))
    if k == 4:
        return "0." + "0" * rng.randrange(0, 30) + str(rng.randrange(1, 10 ** rng.randrange(1, 30)))
    if k == 5:
        return str(rng.randrange(1, 10 ** rng.randrange(1, 40))) + "." + str(rng.randrange(0, 10 ** rng.randrange(1, 30)))
    if k == 6:
        return "%.*e" % (rng.randrange(0, 25), struct.unpack("<d", struct.pack("<Q", rng.getrandbits(62)))[0])
    if k == 7:
        return "-" + str(rng.randrange(0, 10 ** 16))
    return str(2 ** 53 + rng.randrange(-5, 5))


def test_number_parser_matches_strtod():
    rng = random.Random(11)
    undecided = 0
    n = 150000
    for _ in range(n):
        s = _gen(rng)
        if "inf" in s or "nan" in s:
            continue
        r = hostsim.parse_number(s.encode())
        if r is None:
            undecided += 1
            continue
        ref = _libc.strtod(s.encode(), None)
        assert struct.pack("<d", r[0]) == struct.pack("<d", ref), s
    assert undecided < n // 1000


def test_number_parser_hard_cases():
    for s in ["9007199254740993", "2.2250738585072011e-308", "2.2250738585072012e-308", "4.9e-324",
              "2.4703282292062327e-324", "2.4703282292062328e-324", "1.7976931348623157e308",
              "1.7976931348623158e308", "1e400", "1e-400", "0.1", "-0", "123456789012345678901234567890",
              "9007199254740992.5", "7.2057594037927933e16", "0.000000000000000000000000000001"]:
        r = hostsim.parse_number(s.encode())
        assert r is not None, s
        assert struct.pack("<d", r[0]) == struct.pack("<d", _libc.strtod(s.encode(), None)), s
    assert hostsim.parse_number(b"-0")[1] and hostsim.parse_number(b"9007199254740992")[1]
    assert not hostsim.parse_number(b"9007199254740993")[1]  # past 2^53: not exact as an integer
    assert not hostsim.parse_number(b"1.0")[1]


def _nested_filter(kind, n):
    """A filter nesting n comparisons to the right: a op (b op (c op ...))."""
    terms = [f".a == {k + 1}" for k in range(n)]
    if kind == "not":
        terms = [f"(.a == {k + 1} | not)" if k % 3 == 1 else t for k, t in enumerate(terms)]
    op = {"or": "or", "and": "and", "not": "or", "mix": None}[kind]
    q = terms[-1]
    for k in range(n - 2, -1, -1):
        o = op or ("and" if k % 2 else "or")
        q = f"{terms[k]} {o} ({q})"
    return f"select({q}) | count"


@pytest.mark.parametrize("kind", ["or", "and", "not", "mix"])
def test_deep_filters_match_oracle(kind):
    # The RPN of a right-nested filter holds one value per level; past 32 or
    # 64 values a narrow device stack would wrap. Depths 1..65 (64 levels of
    # parentheses, the parser's limit) against the oracle on lines that
    # select at either end of the chain and in the middle.
    lines = [b'{"a":1}', b'{"a":33}', b'{"a":34}', b'{"a":64}', b'{"a":65}', b'{"a":0}', b'{"b":1}', b'{"a":null}']
    for n in (1, 2, 31, 32, 33, 34, 63, 64, 65):
        if kind == "not" and n == 65:
            continue  # a negated term inside the innermost group is a 65th level
        q = _nested_filter(kind, n)
        ev = hostsim.Evaluator(q)
        for ln in lines:
            out, st = oracle.run(q, ln + b"\n", threads=1)
            o, _, _ = ev.eval(ln)
            assert o.status == 0, (n, q, ln)
            assert int(o.selected) == int(out), (kind, n, ln)


def test_rpn_reordering_keeps_every_value():
    # Sethi-Ullman reordering only swaps and/or operands; truth tables of
    # random small filters over three booleans agree with the oracle.
    rng = random.Random(5)

    def gen(d):
        if d == 0 or rng.random() < 0.3:
            return rng.choice([".x", ".y", ".z", ".x == true", ".y == false"])
        k = rng.randrange(3)
        if k == 0:
            return f"({gen(d - 1)} | not)"
        return f"({gen(d - 1)} {'and' if k == 1 else 'or'} {gen(d - 1)})"

    rows = [b'{"x":%s,"y":%s,"z":%s}' % (a, b, c) for a in (b"true", b"false") for b in (b"true", b"false")
            for c in (b"true", b"null")]
    for _ in range(60):
        q = f"select({gen(5)}) | count"
        ev = hostsim.Evaluator(q)
        for ln in rows:
            out, _ = oracle.run(q, ln + b"\n", threads=1)
            assert int(ev.eval(ln)[0].selected) == int(out), (q, ln)


def _hard_numbers(rng):
    """Numbers the 128-bit path tends to leave open: more than 19 digits near
    rounding boundaries, halfway cases between adjacent doubles written out
    in full, subnormals, and digit strings far past 768 digits."""
    from decimal import Decimal, getcontext
    getcontext().prec = 1200
    out = []
    for _ in range(300):
        # exact midpoint of two adjacent doubles (its decimal expansion is finite)
        b = rng.getrandbits(63) & ~(0x7FF << 52) | (rng.randrange(1, 0x7FE) << 52)
        x = struct.unpack("<d", struct.pack("<Q", b))[0]
        y = struct.unpack("<d", struct.pack("<Q", b + 1))[0]
        mid = (Decimal(x) + Decimal(y)) / 2
        s = format(mid, "f") if abs(mid.adjusted()) < 40 else format(mid, "e")
        out.append(s)
        # just above and below the midpoint, in many digits
        d = Decimal(10) ** (mid.adjusted() - 800)
        out.append(format(mid + d, "e"))
        out.append(format(mid - d, "e"))
    for _ in range(300):
        out.append(str(rng.randrange(10 ** 19, 10 ** 40)) + "e" + str(rng.randrange(-360, 300)))
        out.append("0." + "0" * rng.randrange(300, 330) + str(rng.randrange(1, 10 ** 25)))
    out += ["2.4703282292062327208828439643411068618252990130716238221279284125033775363510437593264991818081799618989828234772285886546332835517796989819938739800539093906315035659515570226392290858392449105184435931802849936536152500319370457678249219365623669863658480757001585769269903706311928279558551332927834338409351978015531246597263579574622766465272827220056374006485499977096599470454020828166226237857393450736339007967761930577506740176324673600968951340535537458516661134223766678604162159680461914467291840300530057530849048765391711386591646239524912623653881879636239373280423891018672348497668235089863388587925628302755995657524455507255189313690836254779186948667994968324049705821028513185451396213837722826145437693412532098591327667236328125e-324",
            "2.4703282292062327208828439643411068618252990130716238221279284125033775363510437593264991818081799618989828234772285886546332835517796989819938739800539093906315035659515570226392290858392449105184435931802849936536152500319370457678249219365623669863658480757001585769269903706311928279558551332927834338409351978015531246597263579574622766465272827220056374006485499977096599470454020828166226237857393450736339007967761930577506740176324673600968951340535537458516661134223766678604162159680461914467291840300530057530849048765391711386591646239524912623653881879636239373280423891018672348497668235089863388587925628302755995657524455507255189313690836254779186948667994968324049705821028513185451396213837722826145437693412532098591327667236328125001e-324",
            "1" + "0" * 900, "0." + "0" * 400 + "1", "9" * 850 + "e-600", "1.7976931348623158079e308",
            "1.7976931348623158080e308", "4.9406564584124654e-324", "2.2250738585072013830902327173324040642192159804623318306e-308"]
    return out


def test_exact_decimal_matches_strtod():
    # The wave-wide phase settles every number with the simple decimal
    # conversion; glibc strtod is correctly rounded (the contract of Rust's
    # f64::from_str behind json.rs:73-78). The exact path is checked alone on
    # every case, and behind the fast paths, with nothing left undecided.
    rng = random.Random(23)
    cases = _hard_numbers(rng) + [_gen(rng) for _ in range(20000)]
    cases += ["9007199254740993", "2.2250738585072011e-308", "2.2250738585072012e-308", "4.9e-324",
              "2.4703282292062327e-324", "2.4703282292062328e-324", "1.7976931348623157e308",
              "1.7976931348623158e308", "1e400", "1e-400", "0.1", "-0", "123456789012345678901234567890",
              "9007199254740992.5", "7.2057594037927933e16", "0.000000000000000000000000000001", "0", "-0.0e5"]
    for s in cases:
        if "inf" in s or "nan" in s:
            continue
        b = s.encode()
        ref = struct.pack("<d", _libc.strtod(b, None))
        assert struct.pack("<d", hostsim.slow_decimal(b)) == ref, s
        v = hostsim.parse_number_exact(b)
        assert v is not None and struct.pack("<d", v) == ref, s


def test_deep_lines_walked_with_a_deep_stack():
    # differential.rs:521-533: a 5,000-deep line is valid JSON and the CPU
    # engine decides it; the per-lane walk declines it (64 levels in
    # registers) and the wave-wide phase, whose stack lives in LDS, decides
    # it like the oracle. Mismatched brackets deep down stay invalid.
    deep = b"[" * 5000 + b"]" * 5000
    assert hostsim.validate(deep) == 3
    assert hostsim.validate_deep(deep) == 0
    assert hostsim.validate_deep(b'{"a":' * 3000 + b"1" + b"}" * 3000) == 0
    assert hostsim.validate_deep(b"[" * 5000 + b"]" * 4999 + b"}") == 1
    assert hostsim.validate_deep(b"[" * 4000 + b"{" + b"]" * 4001) == 1
    assert hostsim.validate_deep(b"[" * 5000 + b"]" * 4999) == 1
    assert hostsim.validate_deep(b"[" * 100 + b"]" * 100, cap=100) == 0
    assert hostsim.validate_deep(b"[" * 101 + b"]" * 101, cap=100) == 0  # the innermost empty array takes no level
    assert hostsim.validate_deep(b"[" * 102 + b"]" * 102, cap=100) == 3  # past the LDS stack: declined
    for n in (64, 65, 66, 200):
        line = b'{"k":' + b"[" * n + b'{"x":1}' + b"]" * n + b"}"
        out, st = oracle.run("count", line + b"\n", threads=1)
        assert hostsim.validate_deep(line) == (0 if st["malformed"] == 0 else 1)


def _flat_line(rng):
    keys = ["ts", "host", "status", "bytes", "path", "a", "", "k" * 40, "é", "status", "abcdefghijklmnopq"]
    parts = []
    for _ in range(rng.randrange(0, 9)):
        k = rng.choice(keys)
        v = rng.choice(['"x"', '""', '"GET /a,b:c{}[]"', str(rng.randrange(0, 10 ** rng.randrange(1, 20))), "0",
                        '"%s"' % ("u" * rng.randrange(0, 200)), "12", '"日本"'])
        parts.append('"%s":%s' % (k, v))
    return "{" + ",".join(parts) + "}"


_MUT = ['"', ",", ":", "{", "}", "[", "]", " ", "\t", "\\", "0", "1", "-", ".", "e", "n", "\r", "\x01", "\x7f", "\xff"]


def _mutate(rng, s):
    b = bytearray(s.encode())
    for _ in range(rng.randrange(1, 3)):
        k = rng.randrange(3)
        p = rng.randrange(len(b) + 1)
        c = rng.choice(_MUT).encode("latin-1")
        if k == 0 and p < len(b):
            b[p:p + 1] = c
        elif k == 1:
            b[p:p] = c
        elif p < len(b):
            del b[p]
    return bytes(b).replace(b"\n", b"")


def test_flat_fast_path_never_disagrees_with_the_walk():
    # The bit-parallel fast path may only accept lines the byte walk finds
    # valid, with the same captures, type errors and end; a fuzz of flat
    # lines and one-or-two-byte mutations of them, at every offset within a
    # 16-byte block and with junk after the line. It must also accept the
    # common shape (so the kernels use it).
    rng = random.Random(99)
    qs = ["count", "select(.status == 500) | count", "{h: .host, s: .status, b: .bytes}", ".a.b", ".[0]",
          "select(.path.x == 1) | .k", 'select(.host == "x") | .ts', "group_by(.status) | sum(.bytes)",
          # keys compared as words (empty, UTF-8, several lengths at once) and
          # keys that are not (longer than 16 bytes, holding a quote)
          '{a: .ts, b: .status, c: ."\u00e9", d: .""}', ".abcdefghijklmnopq", '."a\\"b"',
          "{s: .status, l: .abcdefghijklmnopq}"]
    evs = [hostsim.Evaluator(q) for q in qs]
    accepted = 0
    for it in range(12000):
        line = _flat_line(rng).encode()
        if it % 3:
            line = _mutate(rng, line.decode())
        ev = evs[it % len(evs)]
        for lead in (0, 5, 15):
            r = ev.flat_vs_walk(line, lead)
            assert r != 2, (qs[it % len(evs)], line, lead)
            accepted += r == 1
    assert accepted > 12000
    nginx = corpus.nginx(5, 300).split(b"\n")[:-1]
    assert all(evs[1].flat_vs_walk(ln) == 1 for ln in nginx)
