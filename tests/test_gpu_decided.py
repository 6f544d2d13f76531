"""GPU: every line the reference CPU backend decides is decided on the
device, and the runs the host splits or routes specially match the oracle.

* nesting past the 64 levels a lane's walk keeps, numbers the 128-bit path
  leaves open and group keys past 512 KiB go to the wave-wide phase, which
  decides them (LDS deep stack, exact decimal conversion, keys read from
  HBM) instead of declining them (differential.rs:521-533);
* --strict stops at the first malformed or type-error line, as the CPU
  backend does (exec/cpu.rs:401-405);
* a line longer than the chunk runs alone (cli.rs:703-744) and one past the
  line limit is an error (cli.rs:746-760);
* an output expansion past the output buffer (differential.rs:443-518) and
  more group keys than the table holds (cli.rs:839-874) take the host's
  split path;
* the conformance documents (conformance.rs:29-173) through the kernels."""
import json
import math
import os
import random
import struct

import pytest

import corpus
import oracle
import warpjq_amd as w

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def matrix(data, label, chunk=1 << 16, queries=None):
    checked = 0
    for q in queries or corpus.QUERIES:
        for fmt in (w.FMT_NDJSON, w.FMT_CSV, w.FMT_COUNT):
            try:
                cpu, cst = oracle.run(q, data, fmt=fmt, chunk_bytes=chunk, threads=4)
            except oracle.RunError:
                continue
            gpu, gst = w.run(q, data, fmt=fmt, chunk_bytes=chunk)
            assert gpu == cpu, (label, q, fmt)
            for k in ("lines_in", "malformed", "type_errors"):
                assert gst[k] == cst[k], (label, q, fmt, k)
            assert gst["gpu_fallback_lines"] == 0, (label, q)
            checked += 1
    return checked


def test_deeply_nested_lines():
    # differential.rs:521-533, then deep lines with fields on both sides of
    # the deep part, deep objects, and deep lines that are invalid
    n = 5000
    data = b'{"a":1}\n{"a":' + b"[" * n + b"]" * n + b'}\n{"a":2}\n'
    assert matrix(data, "differential deep") > 50
    lines = [b'{"status":500,"host":"pre","bytes":1}']
    for d in (63, 64, 65, 66, 100, 1000, 5000, 20000):
        lines.append(b'{"status":500,"host":"d%d","x":%s,"bytes":%d,"flag":true}' %
                     (d, b"[" * d + b"1" + b"]" * d, d))
        lines.append(b'{"status":404,"n":%s,"host":"o%d","bytes":%d}' % (b'{"k":' * d + b"0" + b"}" * d, d, d))
        lines.append(b'{"status":500,"host":"bad%d","x":%s}' % (d, b"[" * d + b"]" * (d - 1) + b"}"))
    lines.append(b'{"status":200,"host":"post","bytes":2}')
    data = b"\n".join(lines) + b"\n"
    assert matrix(data, "deep mix") > 50
    assert matrix(data, "deep mix, big chunk", chunk=1 << 22) > 50


def test_numbers_the_fast_path_leaves_open():
    # numbers with more than 19 significant digits next to rounding
    # boundaries, exact midpoints, subnormals: compared, summed, grouped
    rng = random.Random(17)
    nums = corpus.hard_numbers(rng, 400)
    lines = []
    for i, s in enumerate(nums):
        lines.append('{"status":%s,"bytes":%s,"host":"h%d","v":"%s"}' % (s, s, i % 7, s))
    data = ("\n".join(lines) + "\n").encode()
    vals = [struct.unpack("<d", struct.pack("<d", float(s)))[0] for s in nums]
    pos = next(s for s in nums if not s.startswith("-"))
    for q in ("count", "min(.bytes)", "max(.bytes)", "select(.bytes > 1e22) | count",
              "select(.bytes <= 2.2250738585072014e-308) | count", "select(.status == %s) | .v" % pos,
              "group_by(.host) | min(.bytes)", "group_by(.host) | max(.bytes)", "group_by(.bytes) | count",
              "select(.bytes >= 1e300) | {b: .bytes, h: .host}", "select(.bytes < 1) | .bytes"):
        for chunk in (1 << 12, 1 << 20):
            cpu, cst = oracle.run(q, data, chunk_bytes=chunk, threads=1)
            gpu, gst = w.run(q, data, chunk_bytes=chunk)
            assert gpu == cpu, (q, chunk)
            assert gst["gpu_fallback_lines"] == 0
    # the sum: the device's is exact to one rounding; the reference adds in
    # line order (agg.rs:62-71), within n * 2^-53 * sum|x| of the exact value
    exact = math.fsum(vals)
    gpu, _ = w.run("sum(.bytes)", data, chunk_bytes=1 << 12)
    cpu, _ = oracle.run("sum(.bytes)", data, chunk_bytes=1 << 12, threads=1)
    g, c = float(gpu), float(cpu)
    bound = len(vals) * 2.0 ** -53 * math.fsum(abs(v) for v in vals)
    assert abs(g - exact) <= 4 * math.ulp(exact) or math.isinf(exact)
    assert abs(c - exact) <= bound


def _strict(q, data, chunk):
    try:
        oracle.run(q, data, on_invalid=oracle.ABORT, chunk_bytes=chunk, threads=4)
        want = None
    except oracle.RunError as e:
        want = (e.kind, e.line)
    try:
        w.run(q, data, on_invalid=w.ABORT, chunk_bytes=chunk)
        got = None
    except w.RunError as e:
        got = (e.code, e.line)
    return want, got


def test_strict_stops_at_type_errors():
    # exec/cpu.rs:401-405: .a.b over {"a":1} is a type error at line 1
    want, got = _strict(".a.b", b'{"a":1}\n', 1 << 16)
    assert want == (2, 1) and got == want
    good = b'{"a":{"b":1},"s":500}\n'
    for te_at, bad_at in ((5, 9), (9, 5), (1, 2), (3000, 100), (100, 3000), (4000, None), (None, 4000)):
        lines = [good] * 5000
        if te_at:
            lines[te_at - 1] = b'{"a":7,"s":500}\n'
        if bad_at:
            lines[bad_at - 1] = b'{"a":\n'
        data = b"".join(lines)
        for q in (".a.b", "select(.a.b == 1) | count", "select(.s == 500) | {x: .a.b}", "sum(.a.b)",
                  "group_by(.s) | sum(.a.b)", "select(.s == 404) | .a.b"):
            for chunk in (1 << 12, 1 << 16, 1 << 22):
                want, got = _strict(q, data, chunk)
                assert got == want, (q, te_at, bad_at, chunk)


def test_strict_across_files():
    files = [b'{"a":{"b":1}}\n' * 300, b'{"a":{"b":2}}\n{"a":3}\n', b'{"a":{"b":1}}\n']
    stream = b"".join(files)
    try:
        oracle.run(".a.b", stream, on_invalid=oracle.ABORT)
        want = None
    except oracle.RunError as e:
        want = e.line
    with pytest.raises(w.RunError) as ei:
        w.run_files(".a.b", files, on_invalid=w.ABORT)
    assert want == 302 and ei.value.line == want


def test_group_keys_past_512_kib():
    keys = ["k" * (600 << 10), "k" * (600 << 10) + "x", "q" * (3 << 20), "short"]
    lines = []
    for i in range(60):
        lines.append('{"host":"%s","bytes":%d}' % (keys[i % 4], i))
    data = ("\n".join(lines) + "\n").encode()
    for q in ("group_by(.host) | count", "group_by(.host) | sum(.bytes)"):
        cpu, _ = oracle.run(q, data, chunk_bytes=16 << 20, threads=2)
        gpu, gst = w.run(q, data, chunk_bytes=16 << 20)
        assert gpu == cpu, q
        assert gst["gpu_fallback_lines"] == 0


def test_line_longer_than_the_chunk():
    # cli.rs:703-744: chunk 1 MiB, line limit 8 MiB, a 2 MiB line, then
    # enough lines for several more chunks
    big = b"x" * (2 << 20)
    data = (b'{"a":1,"tag":"before"}\n' + b'{"a":2,"pad":"' + big + b'"}\n' +
            b"".join(b'{"a":%d,"tag":"after"}\n' % (i % 5) for i in range(200000)))
    for q in ("count", "select(.a == 2) | count", "group_by(.tag) | count", "select(.a == 2) | {a: .a, p: .pad}",
              "select(.a >= 2)", "sum(.a)"):
        cpu, cst = oracle.run(q, data, chunk_bytes=1 << 20, max_line=8 << 20, threads=4)
        gpu, gst = w.run(q, data, chunk_bytes=1 << 20, max_line=8 << 20)
        assert gpu == cpu, q
        assert gst["lines_in"] == cst["lines_in"]
    # cli.rs:746-760: past the limit it is an error, not a fallback
    with pytest.raises(w.RunError) as ei:
        w.run("count", b'{"a":"' + big + b'"}\n', chunk_bytes=1 << 20, max_line=1 << 20)
    assert ei.value.code == 3
    with pytest.raises(oracle.RunError):
        oracle.run("count", b'{"a":"' + big + b'"}\n', chunk_bytes=1 << 20, max_line=1 << 20)


def test_output_expansion_past_the_buffer():
    # differential.rs:443-518 at a real chunk size: 26-byte lines and a
    # six-field projection expand ~6.5x, past the 1.5x + 1 MiB output buffer;
    # the host splits the chunk until the rows fit
    line = b'{"a":1234567890123456789}\n'
    data = line * ((40 << 20) // len(line))
    q = "{alpha: .a, bravo: .a, charlie: .a, delta: .a, echo: .a, foxtrot: .a}"
    cpu, _ = oracle.run(q, data, chunk_bytes=32 << 20, threads=8)
    gpu, gst = w.run(q, data, chunk_bytes=32 << 20)
    assert len(gpu) == len(cpu) and gpu == cpu
    assert gst["gpu_redone_chunks"] > 0
    # differential.rs:443-470: the same shape with tiny chunks
    small = b"".join(b'{"a":%d}\n' % (1000000000 + i) for i in range(40000))
    cpu, _ = oracle.run(q, small, chunk_bytes=1 << 16, threads=4)
    gpu, _ = w.run(q, small, chunk_bytes=1 << 16)
    assert gpu == cpu


def test_more_group_keys_than_the_table():
    # cli.rs:839-874: 200,000 distinct keys. A server with 1 MiB chunks has
    # a 65,536-slot table; a resident dataset goes in one launch, overflows
    # the table and is split by the host until every piece fits.
    data = b"".join(b'{"h":"host-%d","n":%d}\n' % (i, i % 5) for i in range(200000))
    cpu, _ = oracle.run("group_by(.h) | count", data, threads=8)
    srv = w.Server(hbm_budget=1 << 30, workers=1, chunk_bytes=1 << 20)
    try:
        ds = srv.add_dataset(data)
        out, st, _ = srv.wait(srv.submit("group_by(.h) | count", [ds]))
        assert out == cpu
        assert st["gpu_redone_chunks"] > 0
        out, st, _ = srv.wait(srv.submit("group_by(.h) | sum(.n)", [ds]))
        assert out == oracle.run("group_by(.h) | sum(.n)", data, threads=8)[0]
    finally:
        srv.close()


def test_conformance_documents_through_the_kernels():
    # conformance.rs:29-173, one document per line through k_main's LDS
    # arena (and, padded past 1 KiB of their claim, through the wave-wide
    # phase's window reader): accepted ones count as lines in, rejected ones
    # as malformed, exactly as the oracle decides them
    conf = json.load(open(os.path.join(HERE, "golden", "conformance.json"), encoding="utf-8"))
    docs = [d for _, d in conf["must_accept"]] + [d for _, d in conf["must_reject"]]
    docs += [d for _, d, _ in conf["implementation_defined"]]
    docs = [d for d in docs if "\n" not in d]
    data = ("\n".join(docs) + "\n").encode()
    long_pad = ("\n".join(docs) + "\n" + '{"pad":"%s"}\n' % ("p" * 3000)).encode()
    for blob in (data, long_pad):
        for q in (".", "count", "select(.a) | .a", "select(.a != null) | count"):
            cpu, cst = oracle.run(q, blob, threads=1)
            gpu, gst = w.run(q, blob)
            assert gpu == cpu, q
            assert gst["malformed"] == cst["malformed"] and gst["lines_in"] == cst["lines_in"]
    # each document alone, as the only line of its input
    for d in docs:
        b = (d + "\n").encode()
        cpu, cst = oracle.run(".", b, threads=1)
        gpu, gst = w.run(".", b)
        assert (gpu, gst["malformed"]) == (cpu, cst["malformed"]), d


def test_program_tables_past_16_kib():
    # a query whose tables (here a long string literal and a long key) do not
    # fit beside the default line arena: the tables are staged whole and the
    # arena shrinks, the answer is the oracle's
    lit = "x" * 30000
    key = "k" * 9000
    lines = []
    for i in range(3000):
        host = lit if i % 7 == 0 else "h%d" % (i % 5)
        lines.append('{"host":"%s","%s":%d,"bytes":%d}' % (host, key, i, i))
    data = ("\n".join(lines) + "\n").encode()
    for q in ('select(.host == "%s") | count' % lit, 'select(.host != "%s") | sum(.bytes)' % lit,
              'select(.host == "%s") | {b: .bytes, k: .%s}' % (lit, key), "group_by(.%s) | count" % key):
        cpu, _ = oracle.run(q, data, threads=2)
        gpu, gst = w.run(q, data, chunk_bytes=1 << 20)
        assert gpu == cpu, q[:60]
