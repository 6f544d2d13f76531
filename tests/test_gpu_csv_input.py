"""GPU parity of CSV input (warpjq_amd_run_csv): the device transcodes the
records to NDJSON objects and the usual kernels run over them; results equal
the oracle's over its own restatement of the same CSV rules
(oracle.csv_to_ndjson). The reference reads NDJSON only, so this parity is
between two implementations of this build's specification (unpinned against
the reference)."""
import json
import math
import os
import random
import subprocess
import tempfile

import pytest

import oracle
import warpjq_amd as w

pytestmark = pytest.mark.gpu

HEADER = b"host,status,bytes,path,note"


def _field(rng, kind):
    if kind == "num":
        return rng.choice([b"%d" % rng.randrange(0, 10 ** rng.randrange(1, 10)), b"-%d" % rng.randrange(1, 999),
                           b"%d.%d" % (rng.randrange(100), rng.randrange(100)), b"1e%d" % rng.randrange(-5, 20),
                           b"007", b"0", b"1.", b"+3", b""])
    s = rng.choice([b"a", b"web-01", b"x,y", b'say "hi"', b"", b"\xc3\xa9t\xc3\xa9", b"t\tab", b"back\\slash",
                    b"500", b"true", b" lead"])
    if b"," in s or b'"' in s or b"\t" in s or rng.random() < 0.3:
        return b'"' + s.replace(b'"', b'""') + b'"'
    return s


def csv_corpus(seed, rows, bad=0.03):
    rng = random.Random(seed)
    lines = [HEADER]
    for _ in range(rows):
        r = rng.random()
        if r < bad / 3:
            lines.append(b"only,three,fields")
        elif r < 2 * bad / 3:
            lines.append(b'a"b,200,1,/p,n')
        elif r < bad:
            lines.append(b'"unterminated,200,1,/p,n')
        elif r < bad + 0.01:
            lines.append(b"")
        else:
            status = rng.choice([b"200", b"404", b"500", b"503", b'"500"', b""])
            f = [_field(rng, "str"), status, _field(rng, "num"), rng.choice([b"/", b"/api", b'"/a,b"']),
                 _field(rng, "str")]
            lines.append(b",".join(f))
    eol = b"\r\n" if seed % 2 else b"\n"
    return eol.join(lines) + (eol if seed % 3 else b"")


QUERIES = ["count", "select(.status == 500) | count", "select(.status >= 500) | {h: .host, b: .bytes}",
           ".", "select(.note == null) | .path", "max(.bytes)", "min(.bytes)", "group_by(.status) | count",
           "group_by(.path) | max(.bytes)"]


def ulp(x):
    return math.ulp(x) if x else 5e-324


def exact_sums(nd):
    """group key (.host) -> the numeric .bytes values of the NDJSON records."""
    out = {}
    for ln in nd.split(b"\n"):
        if not ln.strip() or ln == b"{":
            continue
        o = json.loads(ln)
        v = o["bytes"]
        e = out.setdefault(o["host"], [])
        if isinstance(v, (int, float)) and not isinstance(v, bool):
            e.append(v)
    return out


def check_sum(k, gpu, cpu, vals):
    # the device sums exactly rounded (i128 and compensated FP64); the
    # oracle, like the reference, in line order (agg.rs:62-71)
    exact = math.fsum(vals)
    mag = math.fsum(abs(v) for v in vals)
    assert abs(gpu - exact) <= 4 * ulp(exact), (k, gpu, exact)
    assert abs(cpu - exact) <= len(vals) * 2.0 ** -53 * mag + 4 * ulp(exact), (k, cpu, exact)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_csv_input_matches_the_oracle(seed):
    data = csv_corpus(seed, 3000)
    for q in QUERIES + ["group_by(.host) | sum(.bytes)", "sum(.bytes)"]:
        for fmt in (w.FMT_NDJSON, w.FMT_COUNT):
            for chunk in (1 << 16, 1 << 20):
                gpu, gst = w.run_csv(q, data, fmt=fmt, chunk_bytes=chunk)
                cpu, cst = oracle.run_csv(q, data, fmt=fmt, chunk_bytes=chunk, threads=4)
                if "sum" not in q:  # sums: checked below
                    assert gpu == cpu, (seed, q, fmt, chunk)
                assert gst["lines_in"] == cst["lines_in"] and gst["malformed"] == cst["malformed"], (seed, q)
    # float sums: within the bounds of the exactly rounded sum
    ex = exact_sums(oracle.csv_to_ndjson(data))
    q = "group_by(.host) | sum(.bytes)"
    gpu = {json.loads(r)["host"]: json.loads(r)["sum"] for r in w.run_csv(q, data)[0].splitlines()}
    cpu = {json.loads(r)["host"]: json.loads(r)["sum"] for r in oracle.run_csv(q, data, threads=4)[0].splitlines()}
    assert set(gpu) == set(cpu) == set(ex)
    for k, vals in ex.items():
        check_sum(k, gpu[k], cpu[k], vals)
    allv = [v for vals in ex.values() for v in vals]
    check_sum(None, float(w.run_csv("sum(.bytes)", data)[0]), float(oracle.run_csv("sum(.bytes)", data)[0]), allv)
    # CSV out of CSV in
    for q in ("select(.status >= 500) | {h: .host, b: .bytes}", "group_by(.host) | count"):
        gpu, _ = w.run_csv(q, data, fmt=w.FMT_CSV)
        cpu, _ = oracle.run_csv(q, data, fmt=w.FMT_CSV, threads=4)
        assert gpu == cpu, q


def test_csv_records_transcode_byte_for_byte():
    # `.` emits every valid record as the device wrote it
    data = csv_corpus(7, 20000, bad=0.0)
    gpu, _ = w.run_csv(".", data)
    nd = oracle.csv_to_ndjson(data)
    assert gpu == b"".join(ln + b"\n" for ln in nd.split(b"\n") if ln)


def test_csv_strict_line_numbers_are_the_files():
    data = b"a,b\n1,2\n\n3,4\n5\n6,7\n"
    with pytest.raises(w.RunError) as e:
        w.run_csv("count", data, on_invalid=w.ABORT)
    assert e.value.line == 5
    assert w.run_csv("count", data)[0] == oracle.run_csv("count", data)[0] == b"3\n"


def test_csv_byte_order_mark():
    data = b"\xef\xbb\xbfhost,status\na,500\nb,200\n"
    for q in ("select(.status == 500) | .host", "."):
        assert w.run_csv(q, data)[0] == oracle.run_csv(q, data)[0]
    assert w.run_csv("select(.status == 500) | .host", data)[0] == b'"a"\n'


def test_csv_header_errors():
    with pytest.raises(w.RunError):
        w.run_csv("count", b"")
    with pytest.raises(w.RunError):
        w.run_csv("count", b"\xef\xbb\xbf")
    with pytest.raises(w.RunError):
        w.run_csv("count", b'"open\n1\n')
    assert w.run_csv("count", b"a,b")[0] == b"0\n"


def test_csv_long_and_expanding_records():
    # groups of 64 records whose bytes pass the LDS stage (long quoted
    # fields), whose output does (control bytes written as \u00XX), and
    # ordinary ones around them
    rng = random.Random(21)
    lines = [b"id,blob,n"]
    for i in range(4000):
        r = rng.random()
        if r < 0.01:
            blob = b'"' + b"x" * rng.randrange(16000, 40000) + b'"'
        elif r < 0.02:
            blob = b'"' + b"\t\x01" * rng.randrange(1000, 3000) + b'"'
        elif r < 0.03:
            blob = b'"' + b'""' * rng.randrange(100, 4000) + b'"'
        else:
            blob = rng.choice([b"a", b'"b,c"', b"", b"12"])
        lines.append(b"%d,%s,%d" % (i, blob, rng.randrange(1000)))
    data = b"\n".join(lines) + b"\n"
    for q in (".", "count", "select(.n > 500) | {id: .id, b: .blob}", "group_by(.blob) | count"):
        gpu, gst = w.run_csv(q, data)
        cpu, cst = oracle.run_csv(q, data, threads=4)
        assert gpu == cpu, q
        assert gst["lines_in"] == cst["lines_in"]


def test_csv_input_at_size():
    # ~9 MiB: many newline tiles and records per transcoding pass
    data = csv_corpus(11, 60000)
    for q in ("select(.status == 500) | count", "group_by(.host) | count", "group_by(.path) | min(.bytes)"):
        gpu, gst = w.run_csv(q, data)
        cpu, cst = oracle.run_csv(q, data, threads=8)
        assert gpu == cpu, q
        assert gst["lines_in"] == cst["lines_in"]


def test_cli_reads_csv_input():
    data = csv_corpus(5, 2000)
    with tempfile.NamedTemporaryFile(suffix=".csv", delete=False) as f:
        f.write(data)
        path = f.name
    try:
        q = "select(.status == 500) | {h: .host}"
        p = subprocess.run([w.CLI_PATH, "--input-csv", q, path], capture_output=True, timeout=120)
        assert p.returncode == 0, p.stderr
        assert p.stdout == oracle.run_csv(q, data, threads=4)[0]
    finally:
        os.unlink(path)
