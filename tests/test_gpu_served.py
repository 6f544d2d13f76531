"""GPU parity of the served form: every request's result in a concurrent
mix equals the same request run alone (warpjq_amd_run_files) and the
oracle's over its datasets read as one stream (chunk.rs:46-57, :198-203);
admission within an HBM budget evicts and reloads without changing results.
"""
import random

import pytest

import corpus
import oracle
import warpjq_amd as w
from test_gpu_batched import BOM, chained

pytestmark = pytest.mark.gpu

TEMPLATES = ["select(.status == {s}) | count", "select(.status >= {s}) | {{h: .host, s: .status}}",
             "group_by(.host) | sum(.bytes)", "sum(.bytes)", "select(.bytes > {b}) | count",
             "group_by(.status) | count", "max(.nested.deep.v)", "select(.status == {s})",
             'select(.host == "{h}") | count', "avg(.bytes)", ".msg", "min(.bytes)"]


def datasets(seed, n):
    rng = random.Random(seed)
    out = []
    for i in range(n):
        k = rng.randrange(3)
        if k == 0:
            out.append(corpus.hostile(seed + i, rng.randrange(20, 400), 0.03))
        elif k == 1:
            out.append(corpus.nginx(seed + i, rng.randrange(50, 2000)))
        else:
            out.append(corpus.hostile(seed + i, rng.randrange(5, 50), 0.0).rstrip(b"\n"))
    return out


def requests(seed, n, n_ds):
    rng = random.Random(seed)
    reqs = []
    for _ in range(n):
        t = TEMPLATES[rng.randrange(len(TEMPLATES))]
        q = t.format(s=rng.choice([200, 404, 500]), b=rng.randrange(50000), h=rng.choice(["a", "b", "c"]))
        ds = rng.sample(range(n_ds), rng.randrange(1, 4))
        fmt = w.FMT_COUNT if rng.random() < 0.2 else w.FMT_NDJSON
        reqs.append((q, ds, fmt))
    return reqs


def test_served_mix_equals_alone():
    data = datasets(5, 12)
    reqs = requests(9, 60, len(data))
    srv = w.Server(hbm_budget=1 << 30, workers=4)
    try:
        ids = [srv.add_dataset(d) for d in data]
        assert ids == list(range(len(data)))
        handles = [srv.submit(q, ds, fmt) for q, ds, fmt in reqs]
        for (q, ds, fmt), h in zip(reqs, handles):
            got, st, lat = srv.wait(h)
            files = [data[i] for i in ds]
            alone, ast = w.run_files(q, files, fmt=fmt)
            assert got == alone, (q, ds)
            cpu, cst = oracle.run(q, chained(files), fmt=fmt, chunk_bytes=1 << 16, threads=4)
            assert got == cpu, (q, ds)
            assert st["lines_in"] == ast["lines_in"] == cst["lines_in"]
            assert lat >= 0
        s = srv.stats()
        assert s["completed"] == len(reqs)
        assert s["p99_ms"] >= s["p50_ms"] > 0
        assert s["evictions"] == 0 and 0 < s["loads"] <= len(data)
    finally:
        srv.close()


def test_served_budget_evicts_and_reloads():
    data = [corpus.nginx(21 + i, 1000 + 37 * i) for i in range(8)]
    largest = max(len(d) for d in data) + 256
    # room for two datasets at a time
    srv = w.Server(hbm_budget=2 * largest + 1024, workers=3)
    try:
        for d in data:
            srv.add_dataset(d)
        q = "group_by(.host) | count"
        order = [i % len(data) for i in range(3 * len(data))]
        handles = [srv.submit(q, [i]) for i in order]
        for i, h in zip(order, handles):
            got, _, _ = srv.wait(h)
            cpu, _ = oracle.run(q, chained([data[i]]))
            assert got == cpu, i
        s = srv.stats()
        assert s["evictions"] > 0
        assert s["reserved"] <= s["budget"]
        # a request bigger than the whole budget fails, and the server goes on
        h = srv.submit("count", list(range(len(data))))
        with pytest.raises(w.RunError):
            srv.wait(h)
        got, _, _ = srv.wait(srv.submit("count", [0]))
        assert got == oracle.run("count", chained([data[0]]))[0]
    finally:
        srv.close()


def test_served_errors_stay_per_request():
    data = [b'{"a":1}\n{"a":}\n', b'{"a":2}\n']
    srv = w.Server(hbm_budget=1 << 24, workers=2)
    try:
        for d in data:
            srv.add_dataset(d)
        bad = srv.submit("count", [0], on_invalid=w.ABORT)
        good = srv.submit("sum(.a)", [0, 1])
        badq = srv.submit("select(", [1])
        with pytest.raises(w.RunError) as e:
            srv.wait(bad)
        assert e.value.line == 2
        assert srv.wait(good)[0] == b"3\n"
        with pytest.raises(w.RunError):
            srv.wait(badq)
    finally:
        srv.close()


def test_reprogram_matches_fresh_context():
    # a context pointed at another query of the same output shape gives the
    # fresh context's result
    data = corpus.nginx(31, 4000)
    buf = w.DeviceBuffer(data)
    chains = [["select(.status == 500) | count", "select(.bytes > 3000) | count", "sum(.bytes)", "max(.bytes)"],
              ["group_by(.host) | count", "group_by(.status) | sum(.bytes)", "group_by(.path) | max(.bytes)"],
              ["select(.status == 404) | {h: .host}", ".path", "select(.status >= 500) | {p: .path, b: .bytes}"]]
    try:
        for chain in chains:
            ctx = w.Context(chain[0], chunk_bytes=1 << 22)
            try:
                for q in chain:
                    ctx.reprogram(q)
                    ctx.submit_device(0, buf.ptr, len(data))
                    r = ctx.wait(0)
                    got = w.result_dict(r)
                    fresh = w.Context(q, chunk_bytes=1 << 22)
                    fresh.submit_device(0, buf.ptr, len(data))
                    want = w.result_dict(fresh.wait(0))
                    fresh.close()
                    got.pop("groups"), want.pop("groups")
                    assert got == want, q
            finally:
                ctx.close()
        ctx = w.Context("count", chunk_bytes=1 << 20)
        with pytest.raises(w.DeviceError):
            ctx.reprogram("group_by(.host) | count")  # other buffers: refused
        ctx.close()
    finally:
        buf.free()


def test_served_rows_count_lines_from_the_resident_index():
    # A small chunk cuts datasets into pieces that start mid-dataset at any
    # byte alignment, so row requests count each claim's lines from the
    # datasets' resident newline indices plus the partial 64-byte blocks at
    # the claim ends. Blank lines, CRLF, a BOM and a missing final newline:
    # rows, line counts and strict-mode line numbers equal the oracle's.
    data = [corpus.nginx(41, 3000), BOM + corpus.hostile(42, 600, 0.02),
            corpus.hostile(43, 300, 0.0).rstrip(b"\n"), b"\n\n" + corpus.nginx(44, 500).replace(b"\n", b"\r\n"),
            b"\n" * 5000 + corpus.nginx(45, 50)]
    srv = w.Server(hbm_budget=1 << 30, workers=3, chunk_bytes=1 << 16)
    try:
        for d in data:
            srv.add_dataset(d)
        qs = ["select(.status >= 500) | {h: .host, s: .status}", "select(.status == 404)", ".msg"]
        sets = [[0], [1, 2], [3, 0, 1], [4, 2], [2, 4, 3, 1, 0]]
        handles = [(q, ds, srv.submit(q, ds)) for q in qs for ds in sets]
        for q, ds, h in handles:
            got, st, _ = srv.wait(h)
            cpu, cst = oracle.run(q, chained([data[i] for i in ds]), chunk_bytes=1 << 16, threads=4)
            assert got == cpu, (q, ds)
            assert st["lines_in"] == cst["lines_in"], (q, ds)
        # the first invalid line's number, past several pieces
        bad = corpus.nginx(46, 2000)
        cut = [i for i, c in enumerate(bad) if c == 10][1500] + 1
        k = srv.add_dataset(bad[:cut] + b'{"status":}\n' + bad[cut:])
        with pytest.raises(w.RunError) as e:
            srv.wait(srv.submit("select(.status == 200)", [0, k], on_invalid=w.ABORT))
        assert e.value.line == data[0].count(b"\n") + 1501 + 1
    finally:
        srv.close()
