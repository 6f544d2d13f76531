"""GPU parity on the nested workload (objects six deep, an escaped key,
1M Zipf-distributed group keys, 30% float values, 0.1% malformed lines).

Counts, minima, maxima, group keys and row outputs are compared with the
oracle byte for byte. Sums and averages of float values depend on summation
order: the oracle adds in line order like the reference (agg.rs:62-71), the
device adds exact integer parts and a compensated float part and rounds
once. Both are checked against the exactly rounded sum (math.fsum over the
values parsed by Python's json): the device within 4 ulp, the oracle within
the recursive-summation bound n * 2^-53 * sum(|x|).
"""
import json
import math

import pytest

import oracle
import warpjq_amd as w

pytestmark = pytest.mark.gpu


def _reject(c):
    raise ValueError(c)


def exact_groups(data, key_path, val_path):
    """key -> (count, numeric, fsum, sum|x|, values) from the lines json accepts."""
    out = {}
    for ln in data.split(b"\n"):
        if not ln.strip():
            continue
        try:
            o = json.loads(ln, parse_constant=_reject)
        except ValueError:
            continue
        k = o if key_path else None
        for p in key_path:
            k = k.get(p) if isinstance(k, dict) else None
        v = o
        for p in val_path:
            v = v.get(p) if isinstance(v, dict) else None
        e = out.setdefault(k, [0, 0, []])
        e[0] += 1
        if isinstance(v, (int, float)) and not isinstance(v, bool):
            e[1] += 1
            e[2].append(v)
    return out


def ulp(x):
    return math.ulp(x) if x else 5e-324


@pytest.fixture(scope="module")
def nested():
    arr, n, lines, bad = w.gen_nested(3 << 20, seed=17)
    return arr[:n].tobytes(), lines, bad


def test_nested_exact_parts(nested):
    data, lines, bad = nested
    for q in ("count", "group_by(.user) | count", "group_by(.user) | min(.bytes)", "group_by(.user) | max(.bytes)",
              "min(.resp.latency)", "max(.bytes)", "select(.resp.code == 500) | group_by(.user) | count",
              ".user", "{u: .user, b: .bytes, f: .resp.meta.d.e.g.f}", "select(.ok) | count",
              'select(.resp.meta.region == "eu-west-2") | {t: .resp.meta.tags}', "group_by(.resp.code) | count"):
        cpu, cst = oracle.run(q, data, chunk_bytes=1 << 20, threads=4)
        gpu, gst = w.run(q, data, chunk_bytes=1 << 20)
        assert gpu == cpu, q
        assert gst["malformed"] == cst["malformed"] == bad
        assert gst["lines_in"] == cst["lines_in"] == lines


@pytest.mark.parametrize("agg", ["sum", "avg"])
@pytest.mark.parametrize("path", [("bytes",), ("resp", "latency")])
def test_nested_float_sums(nested, agg, path):
    data, _, _ = nested
    q = f"group_by(.user) | {agg}(.{'.'.join(path)})"
    ex = exact_groups(data, ("user",), path)
    cpu = {json.loads(r)["user"]: json.loads(r)[agg] for r in oracle.run(q, data, threads=4)[0].splitlines()}
    gpu = {json.loads(r)["user"]: json.loads(r)[agg] for r in w.run(q, data, chunk_bytes=1 << 20)[0].splitlines()}
    assert set(gpu) == set(cpu) == set(ex)
    for k, (cnt, num, vals) in ex.items():
        if num == 0:  # agg.rs: an empty sum is 0, an empty average null
            assert gpu[k] == cpu[k] == (0 if agg == "sum" else None)
            continue
        exact = math.fsum(vals)
        mag = math.fsum(abs(v) for v in vals)
        if agg == "avg":
            exact, mag = exact / num, mag / num
        assert abs(gpu[k] - exact) <= 4 * ulp(exact), (k, gpu[k], exact)
        assert abs(cpu[k] - exact) <= num * 2.0 ** -53 * mag + 4 * ulp(exact), (k, cpu[k], exact)


def test_nested_ungrouped_sum(nested):
    data, _, _ = nested
    ex = exact_groups(data, (), ("bytes",))
    vals = [v for e in ex.values() for v in e[2]]
    gpu = float(w.run("sum(.bytes)", data)[0])
    assert abs(gpu - math.fsum(vals)) <= 4 * ulp(math.fsum(vals))
