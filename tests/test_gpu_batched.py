"""GPU parity of the batched form: many inputs in one launch.

Expected results are the oracle's over the inputs read as one stream, the
way the reference chains a file list (crates/warpjq-core/src/chunk.rs:46-57,
:198-203): each input's BOM is dropped (chunk.rs:210-212, applied per input
by Chunker::new, :224-232) and its last line is complete at its end. Line
numbers continue across inputs, as the reference's GPU backend counts them
(gpu/mod.rs:267, :432).
"""
import random

import pytest

import corpus
import oracle
import warpjq_amd as w

pytestmark = pytest.mark.gpu

BOM = b"\xef\xbb\xbf"


def chained(files):
    out = []
    for f in files:
        if f.startswith(BOM):
            f = f[3:]
        if f and not f.endswith(b"\n"):
            f += b"\n"
        out.append(f)
    return b"".join(out)


def file_mix(seed, n_files, max_lines):
    """Small inputs with the shapes a file list has: empty, blank-only, no
    final newline, a BOM, CRLF, malformed lines."""
    rng = random.Random(seed)
    files = []
    for i in range(n_files):
        k = rng.randrange(6)
        if k == 0:
            f = b""
        elif k == 1:
            f = b"\n" * rng.randrange(1, 3)
        else:
            f = corpus.hostile(seed * 1000 + i, rng.randrange(1, max_lines), 0.05 if k == 2 else 0.0,
                               crlf=(k == 3))
            if k == 4:
                f = f.rstrip(b"\n")  # last line without its newline
            if k == 5:
                f = BOM + f
        files.append(f)
    return files


QUERIES = ["count", "select(.status == 500) | count", ".", "{h: .host, s: .status}", "sum(.bytes)",
           "avg(.bytes)", "min(.bytes)", "max(.nested.deep.v)", "group_by(.host) | count",
           "group_by(.status) | sum(.bytes)", "select(.status >= 404) | {h: .host}", ".arr[1]"]


@pytest.mark.parametrize("chunk", [1 << 14, 1 << 20])
def test_run_files_matches_stream(chunk):
    files = file_mix(7, 300, 40)
    stream = chained(files)
    for q in QUERIES:
        for fmt in (w.FMT_NDJSON, w.FMT_CSV, w.FMT_COUNT):
            try:
                cpu, cst = oracle.run(q, stream, fmt=fmt, chunk_bytes=1 << 16, threads=4)
            except oracle.RunError:
                continue
            gpu, gst = w.run_files(q, files, fmt=fmt, chunk_bytes=chunk)
            assert gpu == cpu, (q, fmt, chunk)
            for k in ("lines_in", "malformed", "type_errors"):
                assert gst[k] == cst[k], (q, fmt, k)
            assert gst["bytes_in"] == sum(len(f) - (3 if f.startswith(BOM) else 0) for f in files)


def test_run_files_large_and_small():
    # inputs larger than the chunk are cut at newlines between small ones
    rng = random.Random(3)
    files = []
    for i in range(40):
        n = rng.choice([1, 3, 20, 900, 2500])
        files.append(corpus.nginx(100 + i, n))
    stream = chained(files)
    for q in ("select(.status == 500) | count", "group_by(.host) | sum(.bytes)",
              "select(.status >= 500) | {p: .path, b: .bytes}"):
        cpu, _ = oracle.run(q, stream, chunk_bytes=1 << 20, threads=4)
        gpu, st = w.run_files(q, files, chunk_bytes=1 << 18)
        assert gpu == cpu, q
        assert st["chunks"] >= 2


def test_run_files_strict_line_numbers():
    files = [b'{"a":1}\n{"a":2}\n', b'{"a":3}', b'\n{"a":}\n{"a":5}\n']
    stream = chained(files)
    with pytest.raises(oracle.RunError) as ce:
        oracle.run("count", stream, on_invalid=oracle.ABORT)
    with pytest.raises(w.RunError) as ge:
        w.run_files("count", files, on_invalid=w.ABORT)
    assert ge.value.line == ce.value.line == 5


def test_run_files_empty_lists():
    for files in ([], [b""], [b"", b""], [BOM], [b"\n"]):
        for q in ("count", "sum(.x)", ".", "group_by(.h) | count"):
            cpu, _ = oracle.run(q, chained(files))
            gpu, _ = w.run_files(q, files)
            assert gpu == cpu, (files, q)


def test_run_files_many_tiny_files():
    # more files than one submit takes (4096 pieces per launch)
    files = [b'{"status":%d,"bytes":%d,"host":"h%d"}' % (500 if i % 7 == 0 else 200, i, i % 13)
             for i in range(9000)]
    stream = chained(files)
    for q in ("count", "select(.status == 500) | count", "group_by(.host) | sum(.bytes)", "max(.bytes)"):
        cpu, _ = oracle.run(q, stream)
        gpu, st = w.run_files(q, files)
        assert gpu == cpu, q
        assert st["chunks"] >= 3


# ---- the device-resident batched submit (warpjq_amd_submit_spans) ---------

def _device(data, offset):
    """`data` in device memory `offset` bytes into a fresh buffer (filler
    before it, 256 readable bytes after it)."""
    t = w.DeviceBuffer(b"#" * offset + data)
    return t, t.ptr + offset


def _result(ctx, slot, r, keys_from, scalar_agg):
    d = w.result_dict(r)
    # group keys by their bytes, merged order-free
    groups = {}
    for i in range(r.n_groups):
        g = r.groups[i]
        key = keys_from(g) if g.key_kind not in (0, 1) else b""  # missing, null
        groups[(g.key_kind, key)] = (g.count, g.numeric, g.min, g.max)
    d["groups"] = groups
    x, gx = ctx.ext(slot)
    d["ext"] = (x.isum_lo, x.isum_hi, x.fsum, x.fcomp) if scalar_agg else None
    d.pop("fallback")
    return d


@pytest.mark.parametrize("q", ["select(.status == 500) | count", "group_by(.host) | sum(.bytes)",
                               "sum(.bytes)", "{h: .host, s: .status}", ".", "group_by(.flag) | count"])
def test_submit_spans_equals_one_stream(q):
    files = [f for f in file_mix(11, 64, 60) if not f.startswith(BOM)]
    files = [f.replace(b"\r", b"") for f in files]
    stream = chained(files)
    keep = []
    spans = []
    for i, f in enumerate(files):
        t, p = _device(f, i % 16)  # every misalignment within a 16-byte block
        keep.append(t)
        spans.append((p, len(f)))
    ts, ps = _device(stream, 0)
    ctx = w.Context(q, chunk_bytes=len(stream) + (1 << 16))
    try:
        ctx.submit_device(0, ps, len(stream))
        r0 = ctx.wait(0)
        agg = q.startswith("sum") or q.endswith("count") and "group_by" not in q
        one = _result(ctx, 0, r0, lambda g: stream[g.key_off:g.key_off + g.key_len], agg)
        ctx.submit_spans(1, spans)
        r1 = ctx.wait(1)
        many = _result(ctx, 1, r1, lambda g: ctx.group_key(1, g.key_off, g.key_len), agg)
        assert many == one
        assert r1.n_lines == stream.count(b"\n")
    finally:
        ctx.close()
    del keep, ts


def test_submit_spans_claim_sizes():
    # results do not depend on the claim size, with spans cut anywhere a
    # line ends
    data = corpus.nginx(5, 3000)
    lines = data.split(b"\n")[:-1]
    rng = random.Random(2)
    files, i = [], 0
    while i < len(lines):
        k = rng.choice([1, 2, 5, 40, 300])
        files.append(b"\n".join(lines[i:i + k]) + (b"\n" if rng.random() < 0.5 else b""))
        i += k
    keep, spans = [], []
    for j, f in enumerate(files):
        t, p = _device(f, (j * 5) % 16)
        keep.append(t)
        spans.append((p, len(f)))
    q = "select(.status == 500) | count"
    want = data.count(b'"status":500')
    for claim in (1024, 4096, 8192):
        ctx = w.Context(q, chunk_bytes=1 << 22)
        try:
            ctx.set_claim(claim)
            ctx.submit_spans(0, spans)
            r = ctx.wait(0)
            assert r.agg.count == want, claim
            assert r.n_lines == len(lines), claim
        finally:
            ctx.close()
    # row outputs: line indices come from the line-count pre-pass, whatever
    # the claim size and however the spans cut the lines
    rq = "select(.status == 500) | {p: .path, b: .bytes}"
    results = []
    for claim in (1024, 3072, 8192):
        ctx = w.Context(rq, chunk_bytes=1 << 22)
        try:
            ctx.set_claim(claim)
            ctx.submit_spans(0, spans)
            r = ctx.wait(0)
            d = w.result_dict(r)
            assert d["n_selected"] == want and d["n_lines"] == len(lines), claim
            results.append((d["out"], d["line_idx"], d["row_off"]))
        finally:
            ctx.close()
    assert results[0] == results[1] == results[2]
    sel = [k for k, ln in enumerate(lines) if b'"status":500' in ln]
    assert results[0][1] == sel
    del keep


def _norm(r, key):
    d = w.result_dict(r)
    gs = []
    for g in d["groups"]:
        g = dict(g)
        g["key"] = key(g["key_off"], g["key_len"]) if g["key_kind"] > 1 else b""
        del g["key_off"]  # where the key bytes sit, not what they are
        gs.append(tuple(sorted(g.items())))
    d["groups"] = sorted(gs)
    return d


@pytest.mark.parametrize("q", ["select(.status == 500) | count", "group_by(.host) | sum(.bytes)",
                               "{h: .host, s: .status}", "sum(.bytes)"])
def test_graph_replays_match_eager(q):
    # captured chains replayed many times, with eager work and syncs between
    # the replays, give the eager result every time
    da, db = corpus.nginx(41, 3000), corpus.hostile(42, 800, 0.05)
    A, B = w.DeviceBuffer(da), w.DeviceBuffer(db)
    eager = w.Context(q, chunk_bytes=1 << 22)
    g = w.Context(q, chunk_bytes=1 << 22)
    g.set_graphs(2)

    def run(ctx, slot, submit):
        submit(ctx, slot)
        if submit is sab:  # a batched submit: keys in the slot's key copy
            return _norm(ctx.wait(slot), lambda o, n: ctx.group_key(slot, o, n))
        data = da if submit is sa else db
        return _norm(ctx.wait(slot), lambda o, n: data[o:o + n])

    sa = lambda c, k: c.submit_device(k, A.ptr, len(da))
    sb = lambda c, k: c.submit_device(k, B.ptr, len(db))
    sab = lambda c, k: c.submit_spans(k, [(A.ptr, len(da)), (B.ptr, len(db))])
    try:
        want_a, want_b = run(eager, 0, sa), run(eager, 0, sb)
        want_ab = run(eager, 1, sab)
        for k in range(5):
            assert run(g, k & 1, sa) == want_a, k
            assert run(eager, 0, sb) == want_b
            assert run(g, k & 1, sb) == want_b, k
            assert run(g, k & 1, sab) == want_ab, k  # three tables, two cached: some captured again
        g.set_claim(4096)  # a new generation: every table captured again
        for k in range(3):
            assert run(g, 0, sa) == want_a
    finally:
        g.close()
        eager.close()
        A.free()
        B.free()


def test_autotune_keeps_results(tmp_path):
    data = corpus.nginx(51, 5000)
    buf = w.DeviceBuffer(data)
    q = "select(.status == 500) | count"
    ctx = w.Context(q, chunk_bytes=1 << 22)
    try:
        ctx.submit_device(0, buf.ptr, len(data))
        want = w.result_dict(ctx.wait(0))
        cache = w.TuneCache(str(tmp_path / "tune.json"))
        claim, arena, ms = ctx.autotune(buf.ptr, len(data), cache)
        assert ms is not None and ms > 0
        assert 1024 <= claim <= 8192 and arena >= claim + 1024
        ctx.submit_device(0, buf.ptr, len(data))
        assert w.result_dict(ctx.wait(0)) == want
        # a second context takes the stored choice without sweeping
        c2 = w.Context(q, chunk_bytes=1 << 22)
        assert c2.autotune(buf.ptr, len(data), w.TuneCache(str(tmp_path / "tune.json"))) == (claim, arena, None)
        c2.submit_device(0, buf.ptr, len(data))
        assert w.result_dict(c2.wait(0)) == want
        c2.close()
    finally:
        ctx.close()
        buf.free()


def test_autotune_group_table_keeps_results(tmp_path):
    # grouped contexts: the autotuner also picks the group table's slots in
    # use; every size gives the same groups, and a table too small for the
    # chunk's keys reports group_overflow rather than a wrong answer
    data = corpus.nginx(52, 8000) + b"".join(b'{"host":"h%d","bytes":%d}\n' % (i, i) for i in range(3000))
    buf = w.DeviceBuffer(data)
    q = "group_by(.host) | sum(.bytes)"
    ctx = w.Context(q, chunk_bytes=1 << 22)
    try:
        ctx.submit_device(0, buf.ptr, len(data))
        r = ctx.wait(0)
        def groups(r):
            return sorted((data[g.key_off:g.key_off + g.key_len], g.count, g.sum)
                          for g in (r.groups[i] for i in range(r.n_groups)))
        want = groups(r)
        claim, arena, ms = ctx.autotune(buf.ptr, len(data), w.TuneCache(str(tmp_path / "t.json")))
        assert ctx.group_table >= 2 * len(want)
        for slots in (4096, 16384, ctx.group_table):
            ctx.set_group_table(slots)
            ctx.submit_device(0, buf.ptr, len(data))
            r = ctx.wait(0)
            assert not r.group_overflow
            assert groups(r) == want
        with pytest.raises(ValueError):
            ctx.set_group_table(3000)
    finally:
        ctx.close()
        buf.free()
