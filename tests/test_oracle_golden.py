"""Pins the oracle (oracle/oracle.cpp) before it is trusted as the checker.

Every assertion restates a known answer from the reference's own tests
(file:line given per test), or a property the mathematics fixes: a textbook
recursive RFC 8259 validator on every short input over a JSON alphabet,
and metamorphic invariants (crates/warpjq-core/tests/invariants.rs).
"""
import itertools
import json
import os

import pytest

import corpus
import oracle as o

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CONF = json.load(open(os.path.join(GOLDEN, "conformance.json"), encoding="utf-8"))

LOG = (b'{"status":200,"bytes":10,"host":"a","msg":"ok"}\n'
       b'{"status":500,"bytes":20,"host":"b","msg":"boom"}\n'
       b'{"status":500,"bytes":30,"host":"a","msg":"boom again"}\n'
       b'{"status":404,"bytes":40,"host":"c"}\n')


def run(q, data=LOG, fmt=o.FMT_NDJSON):
    # cpu.rs:521-532 runs with 64-byte chunks and 4 threads over a cursor
    return o.run(q, data, fmt=fmt, chunk_bytes=64, threads=4, streamed=True)[0]


# ---- conformance.rs:29-173, :235-271, :300-331 ------------------------------

@pytest.mark.parametrize("name,doc", CONF["must_accept"])
def test_must_accept(name, doc):
    assert o.validate(doc.encode())


@pytest.mark.parametrize("name,doc", CONF["must_reject"])
def test_must_reject(name, doc):
    assert not o.validate(doc.encode())


@pytest.mark.parametrize("name,doc,accept", CONF["implementation_defined"])
def test_implementation_defined(name, doc, accept):
    assert o.validate(doc.encode()) == accept


def test_pipeline_verdict_matches_scanner():
    for name, doc in CONF["must_accept"]:
        o.run(".", doc.encode() + b"\n", on_invalid=o.ABORT, chunk_bytes=1 << 16, threads=1)
    for name, doc in CONF["must_reject"]:
        if not doc.strip():
            continue  # a blank line is skipped before parsing
        with pytest.raises(o.RunError):
            o.run(".", doc.encode() + b"\n", on_invalid=o.ABORT, chunk_bytes=1 << 16, threads=1)


# ---- json.rs:715-958 ---------------------------------------------------------

def test_lookup_golden():
    assert o.lookup(b'{"a":1,"b":"x"}', ".a") == ("found", "num", b"1")
    for doc, exp in [(b'{"a":1.0}', b"1.0"), (b'{"a":1e3}', b"1e3"), (b'{"a":-0.0}', b"-0.0"),
                     (b'{"a":123456789012345678901234567890}', b"123456789012345678901234567890")]:
        assert o.lookup(doc, ".a")[2] == exp
    assert o.lookup(b'{"r":{"p":["a","b",{"deep":7}]}}', ".r.p[2].deep") == ("found", "num", b"7")
    assert o.lookup(b'{"a":1}', ".z")[1] == "missing"
    assert o.lookup(b'{"a":1}', ".z.y")[1] == "missing"
    assert o.lookup(b'{"a":[1]}', ".a[9]")[1] == "missing"
    assert o.lookup(b'{"a":1}', ".a.b") == ("type_error",)
    assert o.lookup(b'{"a":{}}', ".a[0]") == ("type_error",)
    assert o.lookup(b'{"a":1,"a":2}', ".a")[2] == b"2"
    assert o.lookup(rb'{"a\"b":5}', '."a\\"b"')[2] == b"5"
    assert o.lookup('{"é":5}'.encode(), '."é"')[2] == b"5"
    assert o.lookup(rb'{"\u00e9":5}', '."é"')[2] == b"5"
    assert o.lookup(rb'{"msg":"}{[],:\"","x":9}', ".x")[2] == b"9"
    assert o.lookup(b'{"a":{"b":[1,2]},"c":0}', ".a") == ("found", "obj", b'{"b":[1,2]}')


def test_truncated_containers_error():
    for doc in ["{", "{ ", "{\t", "[", "[ ", '{"a"', '{"a":', '{"a":1,', '{"a":1,"', "[1,", "[1,2", '{"']:
        assert not o.validate(doc.encode())
        for path in (".", '.""', ".a", ".[0]", ".[4294967295]"):
            o.lookup(doc.encode(), path)  # returns, never crashes
    assert o.lookup(b"{", ".a") == ("invalid",)


def test_rejects_malformed():
    for bad in [b'{"a":}', b'{"a" 1}', b"{a:1}", b'{"a":01}', b'{"a":1.}', b'{"a":1e}', b'{"a":"x}',
                b'{"a":tru}', b'{"a":1}x', b"", b'{"a":"x\ny"}']:
        assert not o.validate(bad)


def test_deep_nesting():
    for n in (1000, 100000, 1000000):
        assert o.validate(b"[" * n + b"]" * n)
        assert not o.validate(b"[" * n)
        assert o.validate(b'{"a":' * n + b"1" + b"}" * n)
    n = 200000
    assert o.lookup(b'{"a":' + b"[" * n + b"]" * n + b',"b":7}', ".b")[2] == b"7"


def test_mismatched_and_awkward():
    for bad in [b"[[[[]]]}", b'{"a":[}', b'{"a":]}', b"[}", b'{"a":1]']:
        assert not o.validate(bad)
    for good in [b"{}", b"[]", b'  {"a" : [ 1 , 2 ] }  ', b'"just a string"', b"null", b"-0",
                 '{"a":"A😀"}'.encode()]:
        assert o.validate(good)


def test_truthiness_and_compare():
    # json.rs:922-958 through select()
    lines = [b'{"a":null}', b'{"a":false}', b'{"b":1}', b'{"a":0}', b'{"a":""}', b'{"a":[]}']
    out = run("select(.a)", b"\n".join(lines) + b"\n")
    assert out == b'{"a":0}\n{"a":""}\n{"a":[]}\n'
    assert o.compare(b'{"a":5}', 'select(.a == "5")') == -1
    assert o.compare(b'{"a":"5"}', "select(.a == 5)") == 1
    assert o.compare(b'{"a":5}', "select(.a == 5)") == 0
    assert o.compare(b'{"a":null}', "select(.a == 0)") == -1
    assert o.compare(b'{"a":null}', "select(.a == null)") == 0
    assert o.compare('{"a":"café"}'.encode(), 'select(.a == "café")') == 0
    assert o.compare(rb'{"a":"caf\u00e9"}', 'select(.a == "café")') == 0
    assert o.compare(b'{"a":9}', "select(.a < 10)") == -1
    assert o.compare(b'{"a":1e3}', "select(.a == 1000)") == 0


def test_unescape():
    assert o.unescape(rb"a\"b\\c\/d\b\f\n\r\t") == b'a"b\\c/d\x08\x0c\n\r\t'
    assert o.unescape(rb"\u0041\u00e9") == "Aé".encode()
    assert o.unescape(rb"\ud83d\ude00") == "😀".encode()
    assert o.unescape(rb"\ud83d") == "�".encode()
    assert o.unescape(rb"\ud83dx") == "�x".encode()
    assert o.unescape(rb"\ude00\ud83d") == "��".encode()
    assert o.unescape(rb"\q") is None


# ---- brute force: a textbook RFC 8259 recursive validator ---------------------

def _textbook_validate(s):
    """Direct recursive-descent transcription of the RFC 8259 grammar, with
    warpjq's whitespace set (space, tab, CR, LF) and no UTF-8 check."""
    n = len(s)

    def ws(i):
        while i < n and s[i] in " \t\r\n":
            i += 1
        return i

    def value(i):
        i = ws(i)
        if i >= n:
            raise ValueError
        c = s[i]
        if c == "{":
            i = ws(i + 1)
            if i < n and s[i] == "}":
                return i + 1
            while True:
                i = ws(i)
                if i >= n or s[i] != '"':
                    raise ValueError
                i = string(i)
                i = ws(i)
                if i >= n or s[i] != ":":
                    raise ValueError
                i = value(i + 1)
                i = ws(i)
                if i < n and s[i] == ",":
                    i += 1
                    continue
                if i < n and s[i] == "}":
                    return i + 1
                raise ValueError
        if c == "[":
            i = ws(i + 1)
            if i < n and s[i] == "]":
                return i + 1
            while True:
                i = value(i)
                i = ws(i)
                if i < n and s[i] == ",":
                    i += 1
                    continue
                if i < n and s[i] == "]":
                    return i + 1
                raise ValueError
        if c == '"':
            return string(i)
        for lit in ("true", "false", "null"):
            if s.startswith(lit, i):
                return i + len(lit)
        return number(i)

    def string(i):
        i += 1
        while i < n:
            c = s[i]
            if c == '"':
                return i + 1
            if c == "\\":
                if i + 1 >= n:
                    raise ValueError
                e = s[i + 1]
                if e in '"\\/bfnrt':
                    i += 2
                elif e == "u" and i + 6 <= n and all(x in "0123456789abcdefABCDEF" for x in s[i + 2:i + 6]):
                    i += 6
                else:
                    raise ValueError
                continue
            if ord(c) < 0x20:
                raise ValueError
            i += 1
        raise ValueError

    def number(i):
        j = i
        if j < n and s[j] == "-":
            j += 1
        if j < n and s[j] == "0":
            j += 1
        elif j < n and s[j] in "123456789":
            while j < n and s[j].isdigit():
                j += 1
        else:
            raise ValueError
        if j < n and s[j] == ".":
            j += 1
            if not (j < n and s[j].isdigit()):
                raise ValueError
            while j < n and s[j].isdigit():
                j += 1
        if j < n and s[j] in "eE":
            j += 1
            if j < n and s[j] in "+-":
                j += 1
            if not (j < n and s[j].isdigit()):
                raise ValueError
            while j < n and s[j].isdigit():
                j += 1
        return j

    try:
        i = value(0)
        return ws(i) == n
    except (ValueError, IndexError):
        return False


def test_brute_force_short_inputs():
    alpha = '{}[]":,01-.e\\ tu'
    count = 0
    for k in range(1, 5):
        for t in itertools.product(alpha, repeat=k):
            s = "".join(t)
            assert o.validate(s.encode()) == _textbook_validate(s), s
            count += 1
    assert count > 60000


def test_brute_force_longer_random():
    import random
    rng = random.Random(5)
    alpha = '{}[]":,0123-.eE+\\ntrufalsenull x\t'
    for _ in range(20000):
        s = "".join(rng.choice(alpha) for _ in range(rng.randrange(1, 14)))
        assert o.validate(s.encode()) == _textbook_validate(s), s
    for line in corpus.hostile(8, 800, 0.2).split(b"\n"):
        if line.strip():
            assert o.validate(line) == _textbook_validate(line.decode()), line


# ---- agg.rs:291-424 -----------------------------------------------------------

def test_format_number_golden():
    assert o.format_number(1.0) == "1"
    assert o.format_number(-0.0) == "0"
    assert o.format_number(1.5) == "1.5"
    assert o.format_number(1e300) == "1e+300"
    assert o.format_number(1e-300) == "1e-300"
    assert o.format_number(0.1 + 0.2) == "0.30000000000000004"
    assert "e" in o.format_number(1e18)
    assert o.format_number(float("inf")) == "1.7976931348623157e+308"
    assert o.format_number(float("nan")) == "null"


def test_format_number_round_trips():
    # shortest round-trip digits: parsing the text gives the same double
    import random
    rng = random.Random(3)
    for _ in range(20000):
        v = rng.uniform(-1e6, 1e6) * 10 ** rng.randrange(-30, 30)
        s = o.format_number(v)
        assert float(s) == v, (v, s)
        assert len(s.replace("-", "").replace("e+", "e")) <= len(repr(v).replace("-", "").replace("e+", "e")) + 2


def test_empty_and_null_aggregates():
    assert run("count", b"") == b"0\n"
    assert run("sum(.x)", b"") == b"0\n"
    assert run("min(.x)", b"") == b"null\n"
    assert run("avg(.x)", b"") == b"null\n"
    data = b'{"x":10}\n{"x":null}\n{"x":20}\n'
    assert run("count", data) == b"3\n"
    assert run("sum(.x)", data) == b"30\n"
    assert run("avg(.x)", data) == b"15\n"
    assert run("min(.x)", data) == b"10\n"


def test_group_order_jq_total_order():
    data = b"".join(b'{"k":%s}\n' % v for v in
                    [b'"b"', b"10", b"null", b'"a"', b"true", b"2", b"false", b"-5"])
    out = run("group_by(.k) | count", data)
    keys = [json.loads(line)["k"] for line in out.splitlines()]
    assert keys == [None, False, True, -5, 2, 10, "a", "b"]


# ---- output.rs:372-481 --------------------------------------------------------

def test_output_golden():
    assert o.json_string(b'a\nb"c\\d\x01') == rb'"a\nb\"c\\d\u0001"'
    assert o.csv_cell(b"plain") == b"plain"
    assert o.csv_cell(b"a,b") == b'"a,b"'
    assert o.csv_cell(b'say "hi"') == b'"say ""hi"""'
    assert run("{n: .n}", b'{"n":1.500}\n') == b'{"n":1.500}\n'
    assert run("{a: .a, b: .b}", b'{"a":"hi","b":null}\n', fmt=o.FMT_CSV) == b"a,b\nhi,\n"
    assert run("{a: .a}", b'{"a":"a\\nb"}\n', fmt=o.FMT_CSV) == b'a\n"a\nb"\n'
    assert run(".", b"{}\n" * 7, fmt=o.FMT_COUNT) == b"7\n"
    assert run("group_by(.req.host) | count", b'{"req":{"host":"a"}}\n') == b'{"host":"a","count":1}\n'


# ---- chunk.rs:475-646 ---------------------------------------------------------

def test_chunks_end_on_newline():
    data = b"aaaa\nbb\ncccccc\ndd\n"
    for size in range(1, len(data) + 3):
        cs = o.chunks(data, size)
        assert sum(c[0] for c in cs) == len(data)
        off = 0
        for i, (ln, _) in enumerate(cs):
            if i + 1 < len(cs):
                assert data[off + ln - 1:off + ln] == b"\n"
            off += ln


def test_chunk_line_numbers_and_long_lines():
    cs = o.chunks(b"1\n2\n3\n4\n5\n", 4)
    expected = 1
    data = b"1\n2\n3\n4\n5\n"
    off = 0
    for ln, first in cs:
        assert first == expected
        expected += data[off:off + ln].count(b"\n")
        off += ln
    long = b"x" * 1000
    data = b"a\n" + long + b"\nb\n"
    cs = o.chunks(data, 16)
    assert sum(c[0] for c in cs) == len(data)
    assert any(c[0] >= 1000 for c in cs)
    assert sum(c[0] for c in o.chunks(b"a\nb", 2)) == 3


def test_lines_semantics():
    assert o.lines(b'{"a":1}\r\n{"b":2}\r\n') == [(1, b'{"a":1}'), (2, b'{"b":2}')]
    assert o.lines(b"a\n\n   \nb\n") == [(1, b"a"), (4, b"b")]
    assert o.lines(b"a\n\x0c\nb") == [(1, b"a"), (3, b"b")]  # 0x0C is ASCII whitespace


def test_bom_and_stream_paths():
    assert run(".", b"\xef\xbb\xbf{}\n") == b"{}\n"
    data = b"".join(b'{"i":%d}\n' % i for i in range(500))
    a = o.run(".i", data, chunk_bytes=97, threads=1, streamed=False)[0]
    b = o.run(".i", data, chunk_bytes=97, threads=1, streamed=True)[0]
    assert a == b and len(a.splitlines()) == 500
    bom = b"\xef\xbb\xbf" + b'{"a":1}\n{"a":2}\n'
    assert o.run(".a", bom, chunk_bytes=8, threads=1, streamed=True)[0] == b"1\n2\n"
    long = b'"' + b"y" * 5000 + b'"'
    data = b'"a"\n' + long + b'\n"z"\n'
    assert o.run(".", data, chunk_bytes=64, threads=1, streamed=True)[0] == data


# ---- cpu.rs:545-707 -----------------------------------------------------------

def test_cpu_golden():
    assert run(".") == LOG
    assert run("select(.status == 500)") == LOG.split(b"\n", 1)[1].rsplit(b"\n", 2)[0] + b"\n"
    assert run("select(.status == 500) | count") == b"2\n"
    assert run("count") == b"4\n"
    assert run("select(.status==404) | {h: .host, s: .status}") == b'{"h":"c","s":404}\n'
    assert run("select(.status==404) | {m: .msg}") == b'{"m":null}\n'
    assert (run("sum(.bytes)"), run("min(.bytes)"), run("max(.bytes)"), run("avg(.bytes)")) == \
        (b"100\n", b"10\n", b"40\n", b"25\n")
    assert run("select(.status==500) | sum(.bytes)") == b"50\n"
    assert run("group_by(.host) | count") == \
        b'{"host":"a","count":2}\n{"host":"b","count":1}\n{"host":"c","count":1}\n'
    assert run("group_by(.host) | sum(.bytes)") == \
        b'{"host":"a","sum":40}\n{"host":"b","sum":20}\n{"host":"c","sum":40}\n'
    assert run("{h: .host, s: .status}", fmt=o.FMT_CSV) == b"h,s\na,200\nb,500\na,500\nc,404\n"
    assert run("select(.status == 500)", fmt=o.FMT_COUNT) == b"2\n"
    data = b"".join(b'{"i":%d}\n' % i for i in range(2000))
    assert run(".i", data).splitlines() == [str(i).encode() for i in range(2000)]
    out, st = o.run(".a", b'{"a":1}\n{not json\n{"a":2}\n', on_invalid=o.SKIP, streamed=True)
    assert out == b"1\n2\n" and st["malformed"] == 1 and st["lines_in"] == 3
    with pytest.raises(o.RunError) as e:
        o.run(".a", b'{"a":1}\n{oops\n', on_invalid=o.ABORT, streamed=True)
    assert e.value.line == 2
    out, st = o.run(".a.b", b'{"a":1}\n{"a":{"b":2}}\n', on_invalid=o.SKIP, streamed=True)
    assert out == b"2\n" and st["type_errors"] == 1
    out = run('select(.status >= 400 and .host == "a")')
    assert b"boom again" in out and len(out.splitlines()) == 1
    assert run("select(.status == 200 or .status == 404) | count") == b"2\n"
    assert run("select(.status == 500 | not) | count") == b"2\n"
    assert run(".", b"") == b""
    assert run("group_by(.h) | count", b"") == b""
    assert run(".a", b'{"a":1}\r\n\r\n{"a":2}\r\n') == b"1\n2\n"


# ---- invariants.rs ------------------------------------------------------------

def test_invariant_chunk_size_and_threads():
    data = corpus.hostile(1, 600, 0.05)
    for q in ("select(.status == 500)", ".host", "group_by(.host) | count", "sum(.bytes)", "count"):
        base = o.run(q, data, chunk_bytes=1 << 20, threads=4)[0]
        for size in (1, 7, 61, 997, 4096, 65536):
            assert o.run(q, data, chunk_bytes=size, threads=4)[0] == base, (q, size)
        for th in (1, 2, 3, 8):
            assert o.run(q, data, chunk_bytes=1 << 16, threads=th)[0] == base, (q, th)


def test_invariant_count_rows_and_group_totals():
    data = corpus.hostile(2, 1500, 0.05)
    rows = o.run("select(.status >= 404)", data)[0].count(b"\n")
    assert o.run("select(.status >= 404) | count", data)[0] == b"%d\n" % rows
    total = int(o.run("count", data)[0])
    groups = [json.loads(line) for line in o.run("group_by(.status) | count", data)[0].splitlines()]
    assert sum(g["count"] for g in groups) == total
    keys = [g["status"] for g in groups]
    assert keys == sorted(set(keys))
    s = int(o.run("sum(.bytes)", data)[0])
    gs = [json.loads(line)["sum"] for line in o.run("group_by(.host) | sum(.bytes)", data)[0].splitlines()]
    assert sum(gs) == s
