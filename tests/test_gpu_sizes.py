"""GPU parity at the workloads' own shapes, beyond the small corpora:

* the 64 MiB nginx workload (the plain method's size) through the four
  queries of the 8 GiB workload plus sum(.bytes), byte for byte;
* 512 MiB of the nested workload (1M Zipf users, 30% float values, 0.1%
  malformed): group membership, counts, minima, maxima and row outputs byte
  for byte; float sums and averages per group within the recursive-summation
  bound of the oracle's line-order sum (agg.rs:62-71), n * 2^-53 * sum|x|,
  with the device's own sum exact to one rounding."""
import json
import math

import pytest

import oracle
import warpjq_amd as w

pytestmark = pytest.mark.gpu

C1_QUERIES = ["select(.status == 500) | count", "select(.status >= 500) | {p: .path, b: .bytes}",
              "group_by(.host) | count", "select(.status == 500)", "sum(.bytes)"]


def test_nginx_64mib_four_queries():
    arr, n, lines, hist = w.gen_nginx(64 << 20, seed=5)
    data = arr[:n].tobytes()
    for q in C1_QUERIES:
        cpu, cst = oracle.run(q, data, threads=16)
        gpu, gst = w.run(q, data)
        assert gpu == cpu, q
        assert gst["lines_in"] == cst["lines_in"] == lines
    assert int(w.run(C1_QUERIES[0], data)[0]) == hist[500]


def _rows(out, key):
    return {json.loads(r)[key]: json.loads(r) for r in out.splitlines()}


def test_nested_512mib():
    arr, n, lines, bad = w.gen_nested(512 << 20, seed=23)
    data = arr[:n].tobytes()
    for q in ("count", "group_by(.user) | count", "group_by(.user) | max(.bytes)", "group_by(.user) | min(.resp.latency)",
              "select(.resp.code == 500) | {u: .user, b: .bytes}"):
        cpu, cst = oracle.run(q, data, threads=16)
        gpu, gst = w.run(q, data)
        assert gpu == cpu, q
        assert gst["malformed"] == cst["malformed"] == bad
    for agg, path in (("sum", "bytes"), ("avg", "resp.latency")):
        q = f"group_by(.user) | {agg}(.{path})"
        cnt = _rows(w.run(f"group_by(.user) | count", data)[0], "user")
        mx = _rows(w.run(f"group_by(.user) | max(.{path})", data)[0], "user")
        mn = _rows(w.run(f"group_by(.user) | min(.{path})", data)[0], "user")
        cpu = _rows(oracle.run(q, data, threads=16)[0], "user")
        gpu = _rows(w.run(q, data)[0], "user")
        assert set(gpu) == set(cpu) == set(cnt)
        worst = 0.0
        for k, g in gpu.items():
            gv, cv = g[agg], cpu[k][agg]
            if gv is None or cv is None:
                assert gv == cv, k
                continue
            m = max(abs(mx[k]["max"] or 0), abs(mn[k]["min"] or 0))
            num = cnt[k]["count"]
            bound = num * 2.0 ** -53 * num * m / (num if agg == "avg" else 1) + 4 * math.ulp(abs(cv) or 5e-324)
            assert abs(gv - cv) <= bound, (k, gv, cv, bound)
            if cv:
                worst = max(worst, abs(gv - cv) / abs(cv))
        print(f"{q}: largest relative difference from the oracle's line-order sum {worst:.3g}")
