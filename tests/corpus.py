"""Seeded NDJSON corpora for the parity tests.

`hostile()` restates the shape of the reference's differential corpus
(crates/warpjq-core/tests/differential.rs:38-234): a predictable outer
object with hostile values inside (unicode, escapes, surrogate pairs,
integers past 2^53, duplicate keys, empty containers, CRLF, blank and
malformed lines). The reference draws from ChaCha8, which is not available
here, so this uses Python's seeded Mersenne Twister over the same value set.

`nginx()` is the benchmark schema (~220 B/line): ts, host (8), method (5),
path (32, Zipf s=1), status (200 80% / 404 10% / 500 5% / 301,304,403,502,503
1% each), bytes (lognormal mu=8 sigma=2, capped at 2^40) and a 40-120 char
printable-ASCII ua without quote or backslash.
"""
import math
import random

KEYS = ["status", "bytes", "host", "msg", "nested", "arr", "flag", "n"]
HOSTS = ["a", "b", "c", "web-01", "ünïcøde", 'with\\"quote', "日本"]

RENORMALISED = [r'"A"', r'"é"', r'"日 mixed with literal text"',
                r'"😀"', r'"\/slash"', '"1e3 is a string here"']

SIMPLE = ["null", "true", "false", "0", "-0", "9007199254740993",
          "123456789012345678901234567890", "1e3", "-2.5E+10", "1.0", "0.1",
          '""', '"plain"', r'"with \"escaped\" quotes"', r'"tab\there\nnewline\\backslash"',
          '"emoji 😀 and é"', '"日本語 ☃ café"', "{}", "[]"]

MALFORMED = ['{"a":}', '{"a" 1}', '{a:1}', '{"a":01}', '{"a":1.}', '{"a":1e}',
             '{"a":"unterminated', '{"a":tru}', '{"a":1}trailing', '[1,2',
             r'{"a":"bad \q escape"}', 'not json at all']


def _q(s):
    """JSON-quote a Python string the way Rust's {:?} would for these keys."""
    out = '"'
    for ch in s:
        if ch == '"':
            out += '\\"'
        elif ch == '\\':
            out += '\\\\'
        else:
            out += ch
    return out + '"'


def hard_value(rng, depth):
    if rng.randrange(14) == 0:
        return RENORMALISED[rng.randrange(len(RENORMALISED))]
    k = rng.randrange(20)
    if k < 19:
        return SIMPLE[k]
    if depth == 0:
        return str(rng.randrange(1000))
    if rng.random() < 0.5:
        n = rng.randrange(4)
        return "[" + ",".join(hard_value(rng, depth - 1) for _ in range(n)) + "]"
    n = rng.randrange(1, 4)
    return "{" + ",".join(f"{_q(KEYS[i % len(KEYS)])}:{hard_value(rng, depth - 1)}" for i in range(n)) + "}"


def good_line(rng):
    status = [200, 301, 404, 500, 503][rng.randrange(5)]
    host = HOSTS[rng.randrange(len(HOSTS))]
    flag = "null" if rng.random() < 0.3 else ("true" if rng.random() < 0.5 else "false")
    return (f'{{"status":{status},"bytes":{rng.randrange(100000)},"host":{_q(host)},'
            f'"msg":{hard_value(rng, 1)},"nested":{{"deep":{{"v":{rng.randrange(500)}}}}},'
            f'"arr":[{hard_value(rng, 0)},{rng.randrange(10)},{hard_value(rng, 0)}],'
            f'"flag":{flag},"n":{hard_value(rng, 2)}}}')


def hostile(seed, lines, bad_ratio=0.0, crlf=False):
    rng = random.Random(seed)
    out = []
    for i in range(lines):
        if i > 0 and i % 37 == 0:
            out.append("\n")
        if i > 0 and i % 53 == 0:
            out.append("   \n")
        out.append(MALFORMED[rng.randrange(len(MALFORMED))] if rng.random() < bad_ratio else good_line(rng))
        out.append("\n")
    s = "".join(out)
    if crlf:
        s = s.replace("\n", "\r\n")
    return s.encode()


# Hand-written edge cases (differential.rs:408-435).
EDGE_LINES = [
    '{}', '{"status":500}', '{"status":500,"status":200}', '{"host":"a","host":"b"}',
    r'{"msg":"}{[],:\" tricky"}', '{"bytes":9007199254740993}', '{"bytes":-0}',
    '{"bytes":1e308}', '{"bytes":1e-308}', r'{"host":"AB"}', r'{"host":"😀"}',
    r'{"host":"\ud83d"}', '{"nested":{"deep":{"v":1}}}', '{"arr":[]}', '{"arr":[[[[1]]]]}',
    '{"flag":false}', '{"flag":null}', '   {"status":404}   ', '{"ünïcøde key":1,"status":500}',
    '{"a":{"b":{"c":{"d":{"e":{"f":{"g":1}}}}}}}',
    # bytes a word-at-a-time special-byte test can flag by borrow right after
    # a real hit ('#' after '"', ']' after a backslash, ' ' after a control)
    '{"host":"#x","status":500}', '{"#k":1,"status":500}', r'{"msg":"a\\]b","status":500}',
    r'{"msg":"q\"#","status":500}', '{"msg":"\t ","status":500}',
] + [
    # the same at every offset within a 16-byte block, and numbers across block edges
    '{"pad":"' + "p" * n + '","host":"#y","n":' + "1" * (1 + n % 17) + ',"status":500}' for n in range(16)
]


def edge_corpus():
    return ("\n".join(EDGE_LINES) + "\n").encode()


# The reference's differential query matrix (differential.rs:250-293).
QUERIES = [
    ".", ".status", ".host", ".msg", ".missing", ".nested.deep.v", ".arr[1]", ".arr[9]",
    "select(.status == 500)", "select(.status != 500)", "select(.status >= 404)",
    "select(.status < 300)", 'select(.host == "a")', 'select(.host == "ünïcøde")',
    "select(.flag)", "select(.flag | not)", "select(.flag == null)", "select(.flag == true)",
    'select(.status == 500 and .host == "b")', "select(.status == 200 or .status == 404)",
    "select(.nested.deep.v > 250)", "{h: .host, s: .status}", "{h: .host, m: .msg, missing: .nope}",
    "{v: .nested.deep.v, a: .arr[0]}", "select(.status == 500) | {h: .host}", "count",
    "select(.status == 500) | count", "sum(.bytes)", "min(.bytes)", "max(.bytes)", "avg(.bytes)",
    "select(.status >= 400) | sum(.bytes)", "sum(.nested.deep.v)", "sum(.missing)",
    "group_by(.host) | count", "group_by(.status) | count", "group_by(.host) | sum(.bytes)",
    "group_by(.flag) | count", "group_by(.status) | avg(.bytes)",
    "select(.status == 500) | group_by(.host) | count",
]

# nginx schema --------------------------------------------------------------

NG_HOSTS = [f"host-{i}.example.com" for i in range(8)]
NG_METHODS = ["GET", "POST", "PUT", "DELETE", "HEAD"]
NG_PATHS = [f"/api/v{1 + i % 3}/{w}" for i, w in enumerate(
    ["users", "orders", "items", "login", "logout", "search", "cart", "checkout", "status",
     "health", "metrics", "images", "static/app.js", "static/app.css", "profile", "settings",
     "billing", "invoices", "reports", "admin", "feed", "comments", "likes", "share", "upload",
     "download", "export", "import", "notifications", "messages", "friends", "groups"])]
_ZIPF_W = [1.0 / (k + 1) for k in range(32)]
_ZIPF_CDF = []
_acc = 0.0
for _w in _ZIPF_W:
    _acc += _w
    _ZIPF_CDF.append(_acc / sum(_ZIPF_W))
UA_ALPHA = "".join(chr(c) for c in range(0x20, 0x7F) if chr(c) not in '"\\')
STATUS_CDF = [(0.80, 200), (0.90, 404), (0.95, 500), (0.96, 301), (0.97, 304), (0.98, 403),
              (0.99, 502), (1.0, 503)]


def nginx_line(rng, i):
    r = rng.random()
    status = next(s for p, s in STATUS_CDF if r < p)
    u = rng.random()
    path = NG_PATHS[next(k for k, c in enumerate(_ZIPF_CDF) if u <= c)]
    b = min(int(math.exp(rng.gauss(8.0, 2.0))), 1 << 40)
    ua = "".join(rng.choice(UA_ALPHA) for _ in range(rng.randrange(40, 121)))
    return (f'{{"ts":{1700000000 + i},"host":"{NG_HOSTS[rng.randrange(8)]}",'
            f'"method":"{NG_METHODS[rng.randrange(5)]}","path":"{path}","status":{status},'
            f'"bytes":{b},"ua":"{ua}"}}')


def nginx(seed, lines):
    rng = random.Random(seed)
    return ("\n".join(nginx_line(rng, i) for i in range(lines)) + "\n").encode()


def hard_numbers(rng, n):
    """JSON numbers a 128-bit decimal-to-binary path tends to leave open:
    exact midpoints between adjacent doubles written in full, values a hair
    above and below them in 20-40 digits, long mantissas with large
    exponents, subnormals."""
    import struct
    from decimal import Decimal, getcontext
    getcontext().prec = 1200
    out = []
    while len(out) < n:
        b = (rng.getrandbits(52)) | (rng.randrange(1, 0x7FE) << 52)
        x = struct.unpack("<d", struct.pack("<Q", b))[0]
        y = struct.unpack("<d", struct.pack("<Q", b + 1))[0]
        mid = (Decimal(x) + Decimal(y)) / 2
        k = rng.randrange(4)
        if k == 0:
            s = format(mid, "e")
        elif k == 1:
            s = format(mid + Decimal(10) ** (mid.adjusted() - 30), ".35e")
        elif k == 2:
            s = format(mid - Decimal(10) ** (mid.adjusted() - 30), ".35e")
        else:
            s = "%de%d" % (rng.randrange(10 ** 20, 10 ** 30), rng.randrange(-340, 280))
        if rng.random() < 0.3:
            s = "-" + s
        out.append(s.replace("E", "e"))
    out.append("4.9406564584124654e-324")
    out.append("2.4703282292062327208828439643411068618252990130716238221279284125033775363510437593264991818081799618989828234772285886546332835517796989819938739800539093906315035659515570226392290858392449105184435931802849936536152500319370457678249219365623669863658480757001585769269903706311928279558551332927834338409351978015531246597263579574622766465272827220056374006485499977096599470454020828166226237857393450736339007967761930577506740176324673600968951340535537458516661134223766678604162159680461914467291840300530057530849048765391711386591646239524912623653881879636239373280423891018672348497668235089863388587925628302755995657524455507255189313690836254779186948667994968324049705821028513185451396213837722826145437693412532098591327667236328125e-324")
    return out
