"""Benchmark: the 8 GiB nginx-schema filter-count on one MI355X.

One step = one pass of `select(.status == 500) | count` over 8 GiB of
nginx-schema NDJSON already resident in HBM (the chunk split in <=4 GiB
pieces, each through warpjq_amd_submit_device). `value` is input GB/s over
the timed steps. Also reported, once each after the timed loop: the other
three queries of the same workload (kernel-only GB/s), the end-to-end
host-buffer run (pinned H2D + kernels), the batched form on the first 1 GiB
cut into 1,024 files (one launch each against one launch for all), the
served form on the same 1 GiB as 8 resident datasets (p50/p99 at 1,000
requests per second), the roofline of the main kernel
(hipEvents around it, on its own stream) and the CPU baseline (the oracle on
a bounded sample, host threads).

`--workload files` measures the batched form instead: 4,096 files of
lognormal sizes clamped to [4 KB, 256 MB], 16 GiB in all, 1% of them holding
a line of up to 1 MB, resident in HBM; the filter-count over all of them
once per file (one launch each, two slots in flight) and batched (spans of
many files per launch, warpjq_amd_submit_spans); then the mix's CSV slice:
`--csv` output of a projection over its first 512 MB of files, resident,
compared byte for byte with the oracle and timed.

`--workload served` measures the served form: 64 nginx datasets of
log-uniform sizes in [1 MB, 2 GB] kept resident by a warpjq_amd server
(200 GB budget, 4 workers), requests from 12 query templates with random
constants arriving as a Poisson process at a swept rate; reports p50/p99
latency (arrival to result) and throughput per rate, and checks that a
sample of requests gives the same bytes alone as in the mix.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
QUERY = "select(.status == 500) | count"
OTHER = ["select(.status >= 500) | {p: .path, b: .bytes}", "group_by(.host) | count", "select(.status == 500)"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def split_points(host, n, piece):
    """Newline-aligned piece boundaries of at most `piece` bytes."""
    pts = [0]
    while n - pts[-1] > piece:
        target = pts[-1] + piece
        j = target - 1
        while host[j] != 10:
            j -= 1
        pts.append(j + 1)
    pts.append(n)
    return list(zip(pts[:-1], pts[1:]))


def file_sizes(rng, n_files, total, lo=4000, hi=256_000_000):
    """Lognormal file sizes clamped to [lo, hi], scaled to sum to `total`."""
    import numpy as np
    s = rng.lognormal(mean=np.log(1 << 20), sigma=2.0, size=n_files)
    for _ in range(50):
        s = np.clip(s * (total / s.sum()), lo, hi)
    return s.astype(np.int64)


def bench_files(args, w, torch, np, oracle):
    n_files = args.files
    total = int(args.size_gib * (1 << 30)) if args.size_gib != 8.0 else 16 << 30
    rng = np.random.default_rng(args.seed)
    t0 = time.time()
    host, n, n_lines, hist = w.gen_nginx(total, seed=args.seed, threads=min(16, os.cpu_count() or 1))
    log(f"generated {n} bytes, {n_lines} lines in {time.time() - t0:.1f}s")
    sizes = file_sizes(rng, n_files, n)
    # newline-aligned file boundaries near the drawn sizes
    bounds, a = [], 0
    for k, sz in enumerate(sizes):
        if a >= n:
            break
        b = n if k == n_files - 1 else min(n, a + int(sz))
        while b < n and host[b - 1] != 10:
            b += 1
        bounds.append((a, b))
        a = b
    if a < n:
        bounds[-1] = (bounds[-1][0], n)
    want = hist[500]
    # 1% of the files get one long line (64 KB - 1 MB) in place of the lines it covers
    n_long = int(round(len(bounds) * args.long_frac))
    long_files = rng.choice(len(bounds), size=n_long, replace=False) if n_long else []
    for f in long_files:
        a, b = bounds[f]
        ln = int(min(b - a, rng.integers(64 << 10, 1 << 20)))
        s0 = a
        e0 = s0 + ln
        while e0 < b and host[e0 - 1] != 10:
            e0 += 1
        if e0 - s0 < 128:
            continue
        region = host[s0:e0].tobytes()
        want -= region.count(b'"status":500,')
        head = b'{"ts":1,"host":"h","method":"GET","path":"/long","status":200,"bytes":1,"ua":"'
        body = head + b"x" * (e0 - s0 - len(head) - 3) + b'"}\n'
        host[s0:e0] = np.frombuffer(body, dtype=np.uint8)
    d = torch.from_numpy(host).to("cuda")
    torch.cuda.synchronize()
    base = d.data_ptr()
    ctx = w.Context(QUERY, chunk_bytes=64 << 20, n_slots=2)

    def per_file():
        count, pend = 0, None
        for i, (a, b) in enumerate(bounds):
            ctx.submit_device(i & 1, base + a, b - a)
            if pend is not None:
                count += ctx.wait(pend).agg.count
            pend = i & 1
        count += ctx.wait(pend).agg.count
        return count

    # batched: as many files per launch as fit 4 GiB of positions (65,536 spans)
    groups, cur, pos = [], [], 0
    claim = 8192
    for a, b in bounds:
        v = ((b - a) + 15 + claim - 1) // claim * claim
        if cur and (pos + v > (4 << 30) - (2 << 20) or len(cur) == 65536):
            groups.append(cur)
            cur, pos = [], 0
        cur.append((base + a, b - a))
        pos += v
    if cur:
        groups.append(cur)

    launch_ms = []

    def batched():
        count, pend = 0, None
        launch_ms.clear()
        for i, g in enumerate(groups):
            ctx.submit_spans(i & 1, g)
            if pend is not None:
                count += ctx.wait(pend).agg.count
                launch_ms.append(round(ctx.timing(pend)[2], 3))
            pend = i & 1
        count += ctx.wait(pend).agg.count
        launch_ms.append(round(ctx.timing(pend)[2], 3))
        return count

    gctx = None
    if args.graphs:
        # the per-file form again, each file's chain captured once and replayed
        gctx = w.Context(QUERY, chunk_bytes=64 << 20, n_slots=2)
        gctx.set_graphs(len(bounds))

    def per_file_graphs():
        count, pend = 0, None
        for i, (a, b) in enumerate(bounds):
            gctx.submit_device(i & 1, base + a, b - a)
            if pend is not None:
                count += gctx.wait(pend).agg.count
            pend = i & 1
        count += gctx.wait(pend).agg.count
        return count

    res = {}
    forms = [("per_file", per_file), ("batched", batched)]
    if gctx is not None:
        forms.append(("per_file_graphs", per_file_graphs))
    for name, fn in forms:
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        best = None
        for _ in range(args.steps):
            t = time.perf_counter()
            c = fn()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
            if c != want:
                raise SystemExit(f"parity: {name} count {c} != generator count {want}")
        res[name] = {"ms": round(best * 1e3, 2), "gbps": round(n / best / 1e9, 1)}
        log(name, res[name])
    speedup = res["per_file"]["ms"] / res["batched"]["ms"]
    del d
    torch.cuda.empty_cache()
    csv = csv_slice(args, w, oracle, host, bounds)
    csv_in = csv_input(args, w, oracle, host, bounds)
    line = {
        "metric": "input GB/s of select(.status == 500) | count over a 4,096-file NDJSON mix resident in HBM "
                  "(batched form)",
        "value": res["batched"]["gbps"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["batched"]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"{len(bounds)} nginx NDJSON files, lognormal sizes clamped to [4 KB, 256 MB], "
                               f"{len(long_files)} with a 64 KB-1 MB line, resident in HBM",
                   "bytes": n, "files": len(bounds), "launches_batched": len(groups),
                   "smallest": int(min(b - a for a, b in bounds)), "largest": int(max(b - a for a, b in bounds)),
                   "grid": ctx.grid},
        "per_file": res["per_file"],
        "batched": res["batched"],
        "speedup_batched_vs_per_file": round(speedup, 2),
        "batched_main_ms_per_launch": launch_ms,
        "per_file_graphs": res.get("per_file_graphs"),
        "csv_slice": csv,
        "csv_input": csv_in,
    }
    print(json.dumps(line), flush=True)


def batched_1g(args, w, np, host, n, base):
    """The batched form on the resident input's first 1 GiB cut into 1,024
    newline-aligned files of lognormal sizes: the filter-count once per file
    (one launch each, two slots in flight) and over all of them in one launch
    (warpjq_amd_submit_spans); both counts must equal one launch over the same
    bytes."""
    cut = min(n, 1 << 30)
    while cut < n and host[cut - 1] != 10:
        cut -= 1
    sizes = file_sizes(np.random.default_rng(args.seed), 1024, cut, lo=4000, hi=64 << 20)
    bounds, a = [], 0
    for k, sz in enumerate(sizes):
        if a >= cut:
            break
        b = cut if k == len(sizes) - 1 else min(cut, a + int(sz))
        while b < cut and host[b - 1] != 10:
            b += 1
        bounds.append((a, b))
        a = b
    if a < cut:
        bounds[-1] = (bounds[-1][0], cut)
    ctx = w.Context(QUERY, chunk_bytes=64 << 20, n_slots=2)
    try:
        ctx.submit_device(0, base, cut)
        want = ctx.wait(0).agg.count

        def per_file():
            count, pend = 0, None
            for i, (a, b) in enumerate(bounds):
                ctx.submit_device(i & 1, base + a, b - a)
                if pend is not None:
                    count += ctx.wait(pend).agg.count
                pend = i & 1
            return count + ctx.wait(pend).agg.count

        def batched():
            ctx.submit_spans(0, [(base + a, b - a) for a, b in bounds])
            return ctx.wait(0).agg.count

        res = {"files": len(bounds), "bytes": cut}
        for name, fn in (("per_file", per_file), ("batched", batched)):
            fn()
            best = None
            for _ in range(3):
                t = time.perf_counter()
                c = fn()
                dt = time.perf_counter() - t
                best = dt if best is None else min(best, dt)
                if c != want:
                    raise SystemExit(f"parity: {name} count {c} != {want} from one launch over the same bytes")
            res[name + "_gbps"] = round(cut / best / 1e9, 1)
        res["speedup"] = round(res["batched_gbps"] / res["per_file_gbps"], 2)
        log("batched 1 GiB", res)
        return res
    finally:
        ctx.close()


def served_1g(args, w, np, host, n):
    """The served form on the input's first 1 GiB cut into 8 datasets of
    log-uniform sizes, kept resident by a server (4 workers): 300 requests
    from the 12 templates arriving as a Poisson process at 1,000 per second;
    p50/p99 latency from arrival to result, and 12 of them again alone on the
    idle server, whose bytes must be the same."""
    import threading
    rng = np.random.default_rng(args.seed + 1)
    cut = min(n, 1 << 30)
    while cut < n and host[cut - 1] != 10:
        cut -= 1
    w8 = np.exp(rng.uniform(np.log(1e6), np.log(5e8), size=8))
    bounds, a = [], 0
    for k, f in enumerate(w8 / w8.sum()):
        b = cut if k == 7 else min(cut, a + int(f * cut))
        while b < cut and host[b - 1] != 10:
            b += 1
        if b > a:
            bounds.append((a, b))
        a = b
    srv = w.Server(hbm_budget=4 << 30, workers=4)
    try:
        ids = [srv.add_dataset(host[a:b]) for a, b in bounds]
        for h in [srv.submit("count", [i]) for i in ids]:  # every dataset resident once
            srv.wait(h)
        reqs = [(served_query(rng), int(rng.integers(len(ids)))) for _ in range(300)]
        gaps = rng.exponential(1.0 / 1000.0, size=len(reqs))
        handles, done = [None] * len(reqs), [None] * len(reqs)

        def collect():
            for k in range(len(reqs)):
                while handles[k] is None:
                    time.sleep(0.0002)
                out, _, lat = srv.wait(handles[k])
                done[k] = (out, lat)

        th = threading.Thread(target=collect)
        th.start()
        t0 = time.perf_counter()
        nxt = t0
        for k, (q, d) in enumerate(reqs):
            nxt += gaps[k]
            dt = nxt - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            handles[k] = srv.submit(q, [ids[d]])
        th.join()
        wall = time.perf_counter() - t0
        for k in rng.choice(len(reqs), size=12, replace=False):
            q, d = reqs[k]
            alone, _, _ = srv.wait(srv.submit(q, [ids[d]]))
            if alone != done[k][0]:
                raise SystemExit(f"parity: served result differs alone vs in the mix: {q} on dataset {d}")
        lats = np.array([x[1] for x in done])
        nbytes = sum(bounds[d][1] - bounds[d][0] for _, d in reqs)
        r = {"datasets": len(ids), "bytes": cut, "requests": len(reqs), "rate_per_s": 1000.0,
             "p50_ms": round(float(np.percentile(lats, 50)), 3), "p99_ms": round(float(np.percentile(lats, 99)), 3),
             "achieved_req_per_s": round(len(reqs) / wall, 1), "gbps": round(nbytes / wall / 1e9, 1)}
        log("served 1 GiB", r)
        return r
    finally:
        srv.close()


CSV_QUERY = "select(.status >= 500) | {host: .host, status: .status, path: .path, bytes: .bytes}"


def slice_files(host, bounds, limit=512_000_000):
    """The mix's first files, about `limit` bytes of them."""
    files, size = [], 0
    for a, b in bounds:
        if size >= limit:
            break
        files.append(host[a:b].tobytes())
        size += b - a
    return files, size


def csv_slice(args, w, oracle, host, bounds):
    """The mix's CSV slice: --csv output of a projection over its first
    512 MB of files, resident (the served form, one request over all of them
    read as one stream), checked byte for byte against the oracle before the
    timed runs."""
    files, size = slice_files(host, bounds)
    srv = w.Server(hbm_budget=4 << 30, workers=1)
    try:
        ids = [srv.add_dataset(f) for f in files]
        out, st, _ = srv.wait(srv.submit(CSV_QUERY, ids, fmt=w.FMT_CSV))
        cpu, _ = oracle.run(CSV_QUERY, b"".join(files), fmt=oracle.FMT_CSV, threads=min(16, os.cpu_count() or 1),
                            chunk_bytes=256 << 20)
        if out != cpu:
            raise SystemExit("parity: the CSV slice differs from the oracle")
        best = None
        for _ in range(max(1, args.steps)):
            t = time.perf_counter()
            out, st, _ = srv.wait(srv.submit(CSV_QUERY, ids, fmt=w.FMT_CSV))
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        r = {"query": CSV_QUERY, "files": len(files), "bytes": size, "rows": out.count(b"\n") - 1,
             "out_bytes": len(out), "ms": round(best * 1e3, 2), "gbps": round(size / best / 1e9, 1),
             "identical_to_oracle": True}
        log("csv slice", r)
        return r
    finally:
        srv.close()


CSV_TABLE = "{ts: .ts, host: .host, method: .method, path: .path, status: .status, bytes: .bytes, ua: .ua}"


def csv_input(args, w, oracle, host, bounds):
    """CSV input: the slice's files as one CSV table (the engine's own --csv
    output of all seven fields), read back by warpjq_amd_run_csv with the
    slice's query (transcoded to NDJSON on the device, then run where it
    lies). Checked byte for byte against the oracle's reading of the same
    table before the timed runs; the same query over the NDJSON the table
    came from is timed beside it (both from host memory, PCIe included)."""
    files, _ = slice_files(host, bounds)
    nd = b"".join(files)
    del files
    table, _ = w.run(CSV_TABLE, nd, fmt=w.FMT_CSV, chunk_bytes=64 << 20)
    out, st = w.run_csv(CSV_QUERY, table, fmt=w.FMT_CSV)
    cpu, _ = oracle.run_csv(CSV_QUERY, table, fmt=oracle.FMT_CSV, threads=min(16, os.cpu_count() or 1),
                            chunk_bytes=256 << 20)
    if out != cpu:
        raise SystemExit("parity: CSV input differs from the oracle")
    src_out, _ = w.run(CSV_QUERY, nd, fmt=w.FMT_CSV)
    best = best_nd = None
    for _ in range(max(1, args.steps)):
        t = time.perf_counter()
        out, st = w.run_csv(CSV_QUERY, table, fmt=w.FMT_CSV)
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
        t = time.perf_counter()
        w.run(CSV_QUERY, nd, fmt=w.FMT_CSV)
        dt = time.perf_counter() - t
        best_nd = dt if best_nd is None else min(best_nd, dt)
    r = {"query": CSV_QUERY, "table": CSV_TABLE, "csv_bytes": len(table), "records": table.count(b"\n") - 1,
         "rows": out.count(b"\n") - 1, "ms": round(best * 1e3, 2), "gbps": round(len(table) / best / 1e9, 2),
         "kernel_ms": round(st["kernel_ms"], 2), "ndjson_bytes": len(nd), "ndjson_ms": round(best_nd * 1e3, 2),
         "identical_to_oracle": True, "rows_equal_ndjson_source": out == src_out}
    log("csv input", r)
    return r


SERVED_HOSTS = ["web-01.prod.us-east-1.example.com", "api-1.prod.eu-west-2.internal",
                "cdn-edge.ap-south-1.example.net", "auth.prod.us-east-1.example.com"]
SERVED_PATHS = ["/", "/api/v1/users", "/search", "/checkout", "/health", "/static/app.js", "/admin"]


def served_query(rng):
    """One of 12 templates with random constants."""
    s = int(rng.choice([200, 404, 500, 301, 503]))
    b = int(rng.integers(100, 100000))
    h = SERVED_HOSTS[int(rng.integers(len(SERVED_HOSTS)))]
    m = ["GET", "POST", "PUT", "DELETE", "HEAD"][int(rng.integers(5))]
    p = SERVED_PATHS[int(rng.integers(len(SERVED_PATHS)))]
    t = [f"select(.status == {s}) | count", f"select(.bytes > {b}) | count", "sum(.bytes)", "avg(.bytes)",
         "max(.bytes)", "group_by(.host) | count", "group_by(.status) | sum(.bytes)",
         f"select(.status == {s}) | group_by(.path) | count", f'select(.method == "{m}") | count',
         f"select(.status >= {s} and .bytes > {b}) | count",
         f'select(.status == 503 and .host == "{h}") | {{p: .path, b: .bytes}}',
         f'select(.path == "{p}") | count']
    return t[int(rng.integers(len(t)))]


def bench_served(args, w, np):
    import threading
    n_ds = args.datasets
    rng = np.random.default_rng(args.seed)
    sizes = np.exp(rng.uniform(np.log(1e6), np.log(2e9), size=n_ds)).astype(np.int64)
    total = int(sizes.sum())
    t0 = time.time()
    host, n, n_lines, hist = w.gen_nginx(total, seed=args.seed, threads=min(16, os.cpu_count() or 1))
    log(f"generated {n} bytes in {time.time() - t0:.1f}s")
    bounds, a = [], 0
    for sz in sizes:
        b = min(n, a + int(sz))
        while b < n and host[b - 1] != 10:
            b += 1
        if b > a:
            bounds.append((a, b))
        a = b
    srv = w.Server(hbm_budget=200 << 30, workers=args.workers, chunk_bytes=args.served_chunk_mib << 20)
    ids = [srv.add_dataset(host[a:b]) for a, b in bounds]
    t0 = time.time()
    for h in [srv.submit("count", [i]) for i in ids]:  # load every dataset once
        srv.wait(h)
    log(f"{len(ids)} datasets resident in {time.time() - t0:.1f}s")
    sizes_b = [b - a for a, b in bounds]
    results = []
    for rate in args.rates:
        reqs = []
        for _ in range(args.requests):
            reqs.append((served_query(rng), int(rng.integers(len(ids)))))
        gaps = rng.exponential(1.0 / rate, size=len(reqs))
        handles = [None] * len(reqs)
        done = [None] * len(reqs)

        def collect():
            for k in range(len(reqs)):
                while handles[k] is None:
                    time.sleep(0.0002)
                out, st, lat = srv.wait(handles[k])
                done[k] = (out, lat)

        th = threading.Thread(target=collect)
        th.start()
        t_start = time.perf_counter()
        nxt = t_start
        for k, (q, d) in enumerate(reqs):
            nxt += gaps[k]
            dt = nxt - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            handles[k] = srv.submit(q, [d])
        th.join()
        wall = time.perf_counter() - t_start
        lats = np.array([x[1] for x in done])
        nbytes = sum(sizes_b[d] for _, d in reqs)
        # the same request alone, on an idle server: identical bytes
        sample = rng.choice(len(reqs), size=min(24, len(reqs)), replace=False)
        for k in sample:
            q, d = reqs[k]
            alone, _, _ = srv.wait(srv.submit(q, [d]))
            if alone != done[k][0]:
                raise SystemExit(f"served result differs alone vs in the mix: {q} on dataset {d}")
        r = {"rate_per_s": rate, "requests": len(reqs), "p50_ms": round(float(np.percentile(lats, 50)), 3),
             "p99_ms": round(float(np.percentile(lats, 99)), 3), "max_ms": round(float(lats.max()), 3),
             "achieved_req_per_s": round(len(reqs) / wall, 1), "gbps": round(nbytes / wall / 1e9, 1)}
        results.append(r)
        log("served", r)
    st = srv.stats()
    srv.close()
    # plain per-request launches beside it: each request alone, one after
    # another, through a fresh context with its dataset copied from host
    # memory (warpjq_amd_run_buffer), as a host without the server would run it
    plain_reqs = [(served_query(rng), int(rng.integers(len(ids)))) for _ in range(args.plain_requests)]
    plain_lat = []
    t_start = time.perf_counter()
    for q, d in plain_reqs:
        a, b = bounds[d]
        t = time.perf_counter()
        w.run(q, host[a:b], chunk_bytes=64 << 20)
        plain_lat.append((time.perf_counter() - t) * 1e3)
    plain_wall = time.perf_counter() - t_start
    plain_bytes = sum(sizes_b[d] for _, d in plain_reqs)
    plain = {"requests": len(plain_reqs), "p50_ms": round(float(np.percentile(plain_lat, 50)), 3),
             "p99_ms": round(float(np.percentile(plain_lat, 99)), 3),
             "req_per_s": round(len(plain_reqs) / plain_wall, 1), "gbps": round(plain_bytes / plain_wall / 1e9, 1)}
    log("plain per-request", plain)
    best = max(results, key=lambda r: r["gbps"])
    line = {
        "metric": "input GB/s of a Poisson mix of 12 query templates over 64 resident nginx datasets (served form)",
        "value": best["gbps"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.requests,
        "warmup": len(ids),
        "ms_per_step": best["p50_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"{len(ids)} nginx datasets, log-uniform sizes in [1 MB, 2 GB], resident in HBM, "
                               f"{args.workers} workers, Poisson arrivals",
                   "bytes_resident": int(sum(sizes_b)), "budget": 200 << 30, "loads": st["loads"],
                   "evictions": st["evictions"], "engines_opened": st["engines_opened"],
                   "engines_reprogrammed": st["engines_reprogrammed"]},
        "rates": results,
        "plain_per_request": plain,
        "served_vs_plain": round(best["gbps"] / plain["gbps"], 2) if plain["gbps"] else None,
    }
    print(json.dumps(line), flush=True)


NESTED_QUERIES = ["group_by(.user) | sum(.bytes)", "group_by(.user) | avg(.resp.latency)",
                  "group_by(.user) | min(.bytes)", "group_by(.user) | max(.bytes)"]


def bench_nested(args, w, torch, np):
    target = int(args.size_gib * (1 << 30)) if args.size_gib != 8.0 else 4 << 30
    t0 = time.time()
    host, n, n_lines, n_bad = w.gen_nested(target, seed=args.seed, threads=min(16, os.cpu_count() or 1))
    log(f"generated {n} bytes, {n_lines} lines ({n_bad} malformed) in {time.time() - t0:.1f}s")
    d = torch.from_numpy(host).to("cuda")
    torch.cuda.synchronize()
    base = d.data_ptr()
    pieces = split_points(host, n, 2 << 30)
    res = {}
    for q in NESTED_QUERIES + ["group_by(.user) | count", "count"]:
        ctx = w.Context(q, chunk_bytes=512 << 20, n_slots=1)
        best, groups, total = None, 0, 0
        for rep in range(args.steps + 1):
            chain, groups, total = 0.0, 0, 0
            for a, b in pieces:
                ctx.submit_device(0, base + a, b - a)
                r = ctx.wait(0)
                if r.group_overflow:
                    raise SystemExit(f"{q}: group table overflow")
                groups += r.n_groups
                total += sum(r.groups[i].count for i in range(r.n_groups)) if r.n_groups else r.agg.count
                chain += ctx.timing(0)[1]
            if rep and (best is None or chain < best):
                best = chain
        ctx.close()
        res[q] = {"kernel_ms": round(best, 2), "gbps": round(n / (best / 1e3) / 1e9, 1), "groups": groups,
                  "lines_counted": total}
        log(q, res[q])
    good = n_lines - n_bad
    # size-independent checks: every well-formed line lands in one group
    if res["count"]["lines_counted"] != good or res["group_by(.user) | count"]["lines_counted"] != good:
        raise SystemExit(f"parity: counted {res['count']['lines_counted']} / "
                         f"{res['group_by(.user) | count']['lines_counted']} lines, generator made {good}")
    g = res[NESTED_QUERIES[0]]
    line = {
        "metric": "input GB/s (kernel chain) of group_by(.user) | sum(.bytes) on 4 GiB nested NDJSON resident in HBM",
        "value": g["gbps"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": g["kernel_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic",
        "config": {"workload": "nested NDJSON (depth 6, escaped keys, 1M Zipf-1.1 users, 30% float values, "
                               "0.1% malformed), resident in HBM", "bytes": n, "lines": n_lines,
                   "malformed": n_bad, "pieces": len(pieces)},
        "queries": res,
    }
    print(json.dumps(line), flush=True)


def parity_count(sample, cut, w):
    """The device's count over the first `cut` bytes of `sample`."""
    out, _ = w.run(QUERY, sample[:cut] if cut < len(sample) else sample, chunk_bytes=64 << 20)
    return int(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size-gib", type=float, default=8.0)
    ap.add_argument("--piece-mib", type=int, default=4095)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--workload", choices=["nginx", "files", "served", "nested"], default="nginx")
    ap.add_argument("--datasets", type=int, default=64)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--requests", type=int, default=600)
    ap.add_argument("--rates", type=float, nargs="+", default=[100.0, 500.0, 2000.0, 5000.0])
    ap.add_argument("--plain-requests", type=int, default=60, help="served: requests run one at a time, unserved")
    ap.add_argument("--served-chunk-mib", type=int, default=64, help="served: device chunk of the server's engines")
    ap.add_argument("--files", type=int, default=4096)
    ap.add_argument("--long-frac", type=float, default=0.01, help="files mix: share of files with a long line")
    ap.add_argument("--graphs", action="store_true", help="files mix: also the per-file form with hipGraph replay")
    ap.add_argument("--no-extras", action="store_true", help="only the timed filter-count loop")
    ap.add_argument("--cpu-sample-mib", type=int, default=1024)
    ap.add_argument("--check-mib", type=int, default=1024, help="bytes compared with the oracle before timing")
    ap.add_argument("--claim-kib", type=int, default=0, help="bytes per wave claim (tuning)")
    ap.add_argument("--arena-kib", type=int, default=0, help="LDS arena per wave (tuning)")
    ap.add_argument("--autotune", action="store_true", help="pick claim / arena with the autotuner first")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy the input once device-to-device (calibrates the HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: runs on one GPU only; several GPUs would each take a share of the chunks "
              "(or files / datasets) with one process per GPU, which is not built", file=sys.stderr)
        sys.exit(2)

    t_start = time.time()
    log("bench: loading numpy, torch and the libraries")
    import numpy as np
    import torch
    import oracle
    import warpjq_amd as w
    log(f"bench: imports done in {time.time() - t_start:.1f}s")

    w.probe()
    dev = w.device_name()
    if args.workload == "files":
        return bench_files(args, w, torch, np, oracle)
    if args.workload == "served":
        return bench_served(args, w, np)
    if args.workload == "nested":
        return bench_nested(args, w, torch, np)
    target = int(args.size_gib * (1 << 30))
    t0 = time.time()
    host, n, n_lines, hist = w.gen_nginx(target, seed=args.seed, threads=min(16, os.cpu_count() or 1))
    log(f"generated {n} bytes, {n_lines} lines in {time.time() - t0:.1f}s on {dev}")

    # Outputs agree before anything is timed (cli/src/bench.rs:179-229): the
    # workload's queries through the product's host driver and through the
    # oracle on the same first --check-mib of the input, byte for byte; a
    # mismatch ends the run with status 3 and no bench line.
    cut = min(n, args.check_mib << 20)
    while cut < n and host[cut - 1] != 10:
        cut -= 1
    sample = host[:cut].tobytes()
    host_threads = min(16, len(os.sched_getaffinity(0)))
    parity = {"sample_bytes": cut, "queries": {}}
    for q in [QUERY] + OTHER + ["sum(.bytes)"]:
        log(f"cmd: warpjq --backend gpu '{q}' <first {cut >> 20} MiB>  vs  warpjq --backend cpu -j {host_threads}")
        t = time.time()
        gpu_out, gst = w.run(q, sample, chunk_bytes=64 << 20)
        t1 = time.time()
        cpu_out, cst = oracle.run(q, sample, threads=host_threads, chunk_bytes=256 << 20)
        log(f"  gpu {t1 - t:.2f}s, cpu {time.time() - t1:.2f}s")
        if gpu_out != cpu_out or gst["lines_in"] != cst["lines_in"] or gst["malformed"] != cst["malformed"]:
            log(f"parity: GPU and oracle differ on {q!r} over the first {cut} bytes")
            raise SystemExit(3)
        rows = gpu_out.count(b"\n")
        parity["queries"][q] = {"rows": rows, "bytes": len(gpu_out), "identical": True}
        if q == "sum(.bytes)":
            # integer field: exact on both sides, relative error 0
            parity["queries"][q]["rel_err_vs_cpu"] = abs(float(gpu_out) - float(cpu_out)) / max(1.0, abs(float(cpu_out)))
    log("parity", parity)

    d = torch.from_numpy(host).to("cuda")  # n + 256 bytes of padding
    torch.cuda.synchronize()
    if args.copy_calibration:
        # one plain copy kernel of known size: reads n + 256 bytes, writes n + 256 bytes
        d2 = torch.empty_like(d)
        d2.copy_(d)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": int(d.numel())}), flush=True)
        return
    base = d.data_ptr()
    pieces = split_points(host, n, args.piece_mib << 20)
    ctx = w.Context(QUERY, chunk_bytes=64 << 20, n_slots=1)
    if args.arena_kib:
        ctx.set_arena(args.arena_kib << 10)
    if args.claim_kib:
        ctx.set_claim(args.claim_kib << 10)
    tuned = None
    if args.autotune:
        c_, a_, ms_ = ctx.autotune(base, pieces[0][1] - pieces[0][0])
        tuned = {"claim": c_, "arena": a_, "sample_main_ms": round(ms_, 3)}
        log("autotune", tuned)

    def step():
        # one stream: pieces run back to back, so per-kernel events are clean
        total = 0
        main_ms = 0.0
        for a, b in pieces:
            ctx.submit_device(0, base + a, b - a)
            r = ctx.wait(0)
            total += r.agg.count
            main_ms += ctx.timing(0)[2]
        return total, main_ms

    for _ in range(args.warmup):
        step()

    torch.cuda.synchronize()
    t = time.perf_counter()
    mains = []
    for _ in range(args.steps):
        count, m = step()
        mains.append(m)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    if count != hist[500]:  # a size-independent check: the generator counted its own 500s
        log(f"parity: device count {count} != generator count {hist[500]}")
        raise SystemExit(3)
    ms_per_step = wall * 1e3 / args.steps
    gbps = n / (ms_per_step / 1e3) / 1e9
    main_ms = sum(mains) / len(mains)
    achieved = n / (main_ms / 1e3) / 1e9  # algorithmic bytes (input) / main-kernel time, per step

    extras = {}
    if not args.no_extras:
        # the rest of the workload's queries, kernel-only, one pass each
        for q in OTHER:
            c2 = w.Context(q, chunk_bytes=512 << 20, n_slots=1)
            ps = split_points(host, n, 512 << 20)
            chain = 0.0
            rows = 0
            out_bytes = 0
            for a, b in ps:
                c2.submit_device(0, base + a, b - a)
                r = c2.wait(0)
                rows += r.n_selected + r.n_groups
                out_bytes += r.out_len
                chain += c2.timing(0)[1]
            c2.close()
            extras[q] = {"kernel_gbps": round(n / (chain / 1e3) / 1e9, 1), "rows": rows,
                         "alg_gbps": round((n + out_bytes) / (chain / 1e3) / 1e9, 1)}
            log(q, extras[q])
        # end to end from a host buffer (pinned staging + H2D + kernels + D2H):
        # the first call in the process also starts the staging threads and
        # pins their buffers; the second is the steady state reported
        e2e_runs = []
        for _ in range(2):
            t = time.perf_counter()
            out, st = w.run(QUERY, host, chunk_bytes=64 << 20, n=n)
            e2e_runs.append(time.perf_counter() - t)
            if int(out) != hist[500]:
                raise SystemExit("parity: the end-to-end count differs from the generator's")
        e2e = e2e_runs[-1]
        extras["e2e"] = {"gbps": round(n / e2e / 1e9, 2), "gbps_first_call": round(n / e2e_runs[0] / 1e9, 2),
                         "h2d_ms": round(st["h2d_ms"], 1),
                         "kernel_ms": round(st["kernel_ms"], 1), "chunks": st["chunks"],
                         "h2d_gbps": round(n / (st["h2d_ms"] / 1e3) / 1e9, 1) if st["h2d_ms"] else None}
        log("e2e", extras["e2e"])
        extras["batched_1g"] = batched_1g(args, w, np, host, n, base)
        extras["served_1g"] = served_1g(args, w, np, host, n)
        # the nested workload's group_by over 1M users on a 1 GiB sample
        # (kernel chain; the full 4 GiB form is --workload nested)
        nh, nn, nlines, nbad = w.gen_nested(1 << 30, seed=args.seed, threads=min(16, os.cpu_count() or 1))
        nd = torch.from_numpy(nh).to("cuda")
        torch.cuda.synchronize()
        for q in ("group_by(.user) | sum(.bytes)", "count"):
            c3 = w.Context(q, chunk_bytes=512 << 20, n_slots=1)
            chain, groups = 0.0, 0
            for rep in range(2):
                chain, groups = 0.0, 0
                for a, b in split_points(nh, nn, 1 << 30):
                    c3.submit_device(0, nd.data_ptr() + a, b - a)
                    r = c3.wait(0)
                    if r.group_overflow:
                        raise SystemExit(f"{q}: group table overflow")
                    groups += r.n_groups
                    chain += c3.timing(0)[1]
            c3.close()
            extras["nested_1g " + q] = {"kernel_gbps": round(nn / (chain / 1e3) / 1e9, 1), "groups": groups}
            log("nested", q, extras["nested_1g " + q])
        del nd

    # CPU baseline: the oracle (the reference CPU backend restated) on the
    # parity sample, at -j 8 (the reference's published 8-core figure,
    # docs/ARCHITECTURE.md:268) and at every host core this process may use;
    # best of 3 each. Informative only: it is never the target.
    cpu = None
    cpu_cut = min(cut, args.cpu_sample_mib << 20)
    while cpu_cut < cut and sample[cpu_cut - 1] != 10:
        cpu_cut -= 1
    cpu_sample = sample[:cpu_cut] if cpu_cut < cut else sample
    legs = {}
    for j in sorted({8, host_threads}):
        log(f"cmd: warpjq --backend cpu -j {j} '{QUERY}' <first {cpu_cut >> 20} MiB>")
        best = None
        for _ in range(3):
            t = time.perf_counter()
            out, _ = oracle.run(QUERY, cpu_sample, threads=j, chunk_bytes=256 << 20)
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        if int(out) != int(parity_count(sample, cpu_cut, w)):
            log("cpu baseline: count differs from the device on the baseline sample")
            raise SystemExit(3)
        legs[j] = round(cpu_cut / best / 1e9, 3)
    cpu = {"value": legs[host_threads], "unit": "GB/s", "cores": host_threads, "kind": "port",
           "value_j8": legs.get(8), "cores_visible": os.cpu_count(),
           "sample": f"first {cpu_cut >> 20} MiB ({cpu_cut} B) of the same input, {QUERY}, best of 3, "
                     f"-j {host_threads} (value) and -j 8 (value_j8)"}

    # HBM bytes per k_main launch from the committed rocprofv3 counter runs
    # (FETCH_SIZE and WRITE_SIZE per input byte, calibrated on a plain copy)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_main_nginx_traffic.txt")
    if os.path.exists(tpath):
        kv = {}
        for ln in open(tpath):
            if ":" in ln and not ln.startswith("#"):
                k, v = ln.split(":", 1)
                kv[k.strip()] = v.strip()
        try:
            traffic = round((float(kv["fetch_bytes_per_input_byte"]) + float(kv["write_bytes_per_input_byte"])) * n)
        except (KeyError, ValueError):
            traffic = None

    line = {
        "metric": "input GB/s (kernel) of select(.status == 500) | count on 8 GiB nginx NDJSON resident in HBM",
        "value": round(gbps, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "nginx NDJSON, select(.status == 500) | count, resident in HBM",
                   "bytes": n, "lines": n_lines, "avg_line_bytes": round(n / n_lines, 1),
                   "pieces": len(pieces), "device": dev, "grid": ctx.grid, "autotuned": tuned},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "kernel": "k_main", "main_ms_per_step": round(main_ms, 3)},
        "cpu_baseline": cpu,
        "parity": parity,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
