"""GPU parity: the HIP path (through the C ABI and the host driver) against
the oracle on the same bytes. Bit-exact output bytes, counts and stats."""
import pytest

import corpus
import oracle
import warpjq_amd as w

pytestmark = pytest.mark.gpu

LOG = (b'{"status":200,"bytes":10,"host":"a","msg":"ok"}\n'
       b'{"status":500,"bytes":20,"host":"b","msg":"boom"}\n'
       b'{"status":500,"bytes":30,"host":"a","msg":"boom again"}\n'
       b'{"status":404,"bytes":40,"host":"c"}\n')


def both(q, data, fmt=w.FMT_NDJSON, chunk=1 << 16, on_invalid=w.SKIP):
    cpu, cst = oracle.run(q, data, fmt=fmt, on_invalid=on_invalid, chunk_bytes=chunk, threads=4)
    gpu, gst = w.run(q, data, fmt=fmt, on_invalid=on_invalid, chunk_bytes=chunk)
    return cpu, cst, gpu, gst


def test_probe_and_device():
    w.probe()
    assert "gfx950" in w.device_name()


@pytest.mark.parametrize("q,expect", [
    (".", LOG),
    ("select(.status == 500) | count", b"2\n"),
    ("count", b"4\n"),
    ("select(.status==404) | {h: .host, s: .status}", b'{"h":"c","s":404}\n'),
    ("select(.status==404) | {m: .msg}", b'{"m":null}\n'),
    ("sum(.bytes)", b"100\n"), ("min(.bytes)", b"10\n"), ("max(.bytes)", b"40\n"), ("avg(.bytes)", b"25\n"),
    ("select(.status==500) | sum(.bytes)", b"50\n"),
    ("group_by(.host) | count",
     b'{"host":"a","count":2}\n{"host":"b","count":1}\n{"host":"c","count":1}\n'),
    ("group_by(.host) | sum(.bytes)",
     b'{"host":"a","sum":40}\n{"host":"b","sum":20}\n{"host":"c","sum":40}\n'),
])
def test_log_golden(q, expect):
    # crates/warpjq-core/src/exec/cpu.rs:534-605
    out, _ = w.run(q, LOG)
    assert out == expect


def test_log_csv_and_count_only():
    assert w.run("{h: .host, s: .status}", LOG, fmt=w.FMT_CSV)[0] == b"h,s\na,200\nb,500\na,500\nc,404\n"
    assert w.run("select(.status == 500)", LOG, fmt=w.FMT_COUNT)[0] == b"2\n"


@pytest.mark.parametrize("label,data", ids=lambda v: v if isinstance(v, str) else "data", argvalues=[
    ("clean", corpus.hostile(1, 4000, 0.0)),
    ("15% malformed", corpus.hostile(2, 4000, 0.15)),
    ("crlf", corpus.hostile(3, 2000, 0.05, crlf=True)),
    ("edge", corpus.edge_corpus()),
    ("all malformed", corpus.hostile(4, 500, 1.0)),
])
def test_hostile_matrix(label, data):
    # differential.rs:340-435: every query x format, GPU == CPU bytes
    checked = 0
    for q in corpus.QUERIES:
        for fmt in (w.FMT_NDJSON, w.FMT_CSV, w.FMT_COUNT):
            try:
                cpu, cst = oracle.run(q, data, fmt=fmt, chunk_bytes=1 << 16, threads=4)
            except oracle.RunError:
                continue  # --csv on a whole-line query: rejected by both
            gpu, gst = w.run(q, data, fmt=fmt, chunk_bytes=1 << 16)
            assert gpu == cpu, (label, q, fmt)
            for k in ("lines_in", "malformed", "type_errors"):
                assert gst[k] == cst[k], (label, q, fmt, k)
            checked += 1
    assert checked > 80


def test_empty_and_tiny():
    for data in (b"", b"\n", b"{}\n", b"{}", b"\n\n\n"):
        for q in (".", "count", "sum(.x)", "min(.x)", "group_by(.h) | count"):
            cpu, _, gpu, _ = both(q, data)
            assert gpu == cpu, (data, q)


def test_nginx_filter_count():
    data = corpus.nginx(11, 20000)
    for q in ("select(.status == 500) | count", "select(.status >= 500) | {p: .path, b: .bytes}",
              "group_by(.host) | count", "select(.status == 500)", "sum(.bytes)", "avg(.bytes)"):
        cpu, cst, gpu, gst = both(q, data, chunk=1 << 20)
        assert gpu == cpu, q
        assert gst["lines_in"] == cst["lines_in"]


def test_flat_lines_after_runs_of_nested_ones():
    # A wave that sees a round with no flat line walks the next rounds byte by
    # byte; flat lines arriving then (and long runs of either kind, in
    # several chunks) give the oracle's result whichever path takes them.
    nested = corpus.hostile(12, 3000, 0.02).split(b"\n")[:-1]
    flat = corpus.nginx(13, 9000).split(b"\n")[:-1]
    lines = []
    for k in range(6):
        lines += nested[500 * k:500 * (k + 1)]
        lines += flat[1500 * k:1500 * (k + 1)]
    data = b"\n".join(lines) + b"\n"
    for q in ("select(.status == 500) | count", "group_by(.host) | count", "select(.status >= 500) | {h: .host, s: .status}",
              "sum(.bytes)", "count"):
        for chunk in (1 << 16, 1 << 20):
            cpu, cst, gpu, gst = both(q, data, chunk=chunk)
            assert gpu == cpu, (q, chunk)
            assert gst["lines_in"] == cst["lines_in"] and gst["malformed"] == cst["malformed"], (q, chunk)


def long_lines_corpus(seed):
    """Short lines around lines of 9 KB - 1.2 MB: long strings, long arrays,
    deep objects, a long invalid line and a long line whose key is late."""
    import random
    rng = random.Random(seed)
    short = corpus.nginx(seed, 400).split(b"\n")[:-1]
    longs = []
    for n in (9000, 20000, 70000, 300000, 1200000):
        longs.append(b'{"status":500,"bytes":%d,"host":"h-long","ua":"%s"}' % (n, b"x" * n))
        longs.append(b'{"ua":"%s","status":404,"host":"late","bytes":7}' % (b"y\\\"z" * (n // 4)))
        longs.append(b'{"arr":[%s],"status":200,"bytes":1,"host":"arr"}' % b",".join(b"%d" % i for i in range(n // 8)))
        longs.append(b'{"status":500,"host":"bad","ua":"%s' % (b"w" * n))  # unterminated string: invalid
    deep = b'{"n":' * 50 + b"1" + b"}" * 50
    longs.append(b'{"status":500,"host":"deep","pad":"%s","d":%s}' % (b"p" * 20000, deep))
    lines = short[:]
    for ln in longs:
        lines.insert(rng.randrange(len(lines) + 1), ln)
    return b"\n".join(lines) + b"\n"


def test_long_lines():
    data = long_lines_corpus(4)
    for q in ("count", "select(.status == 500) | count", "sum(.bytes)", "max(.bytes)",
              "group_by(.host) | count", "group_by(.host) | sum(.bytes)", "select(.status == 404) | {h: .host}",
              ".host", "select(.bytes > 100000) | .bytes"):
        for fmt in (w.FMT_NDJSON, w.FMT_COUNT):
            cpu, cst, gpu, gst = both(q, data, fmt=fmt, chunk=1 << 22)
            assert gpu == cpu, (q, fmt)
            for k in ("lines_in", "malformed", "type_errors"):
                assert gst[k] == cst[k], (q, k)


def test_long_line_group_key_across_window():
    # long lines walked by the whole wave: a group key that crosses the LDS
    # window edge (18 KiB) and keys of every length around 32 bytes
    lines = []
    for i, pad in enumerate((18000, 18400, 18430, 36000, 50000)):
        key = ("k%d-" % (i % 2)) + "z" * 1000
        lines.append('{"pad":"%s","host":"%s","bytes":%d}' % ("p" * pad, key, i))
    for n in range(28, 37):
        lines.append('{"host":"%s","bytes":1}' % ("h" * n))
        lines.append('{"host":"%s","bytes":2}' % ("h" * n))
    data = ("\n".join(lines * 3) + "\n").encode()
    for q in ("group_by(.host) | count", "group_by(.host) | sum(.bytes)"):
        cpu, cst, gpu, gst = both(q, data, chunk=1 << 20)
        assert gpu == cpu, q


def test_long_line_string_ends_at_every_offset():
    # a long line's string whose closing quote falls at each offset of the
    # 1 KiB steps of the wave-wide scan and around the LDS window edge
    for kl in (5, 40, 1000):
        lines = []
        for pad in list(range(18416, 18450)) + list(range(17392, 17420)) + [36860, 36870, 9000, 10240]:
            key = "k" + "z" * (kl - 1)
            lines.append('{"pad":"%s","host":"%s","bytes":%d}' % ("p" * pad, key, pad))
        data = ("\n".join(lines) + "\n").encode()
        for q in ("group_by(.host) | count", "sum(.bytes)", ".bytes", 'select(.host == "k") | count'):
            cpu, cst, gpu, gst = both(q, data, chunk=1 << 22)
            assert gpu == cpu, (kl, q)
            assert gst["malformed"] == cst["malformed"] == 0
