"""GPU parity of inputs named by path and of the command line.

warpjq_amd_run_paths opens files the way the reference's CLI does
(Input::open / Input::stdin / Input::chain, crates/warpjq-core/src/chunk.rs:
60-205): regular files mapped, stdin and pipes read as a stream with the
partial last line carried into the next read (stream_chunks, chunk.rs:292-
387), the whole list one stream. Expected results are the oracle's over the
same bytes read as one stream, as in test_gpu_batched.py. The command line's
stdout, stderr summary and exit codes follow crates/warpjq-cli/src/run.rs:
16-150 and main.rs:155-190.
"""
import os
import subprocess
import threading

import pytest

import corpus
import oracle
import warpjq_amd as w
from test_gpu_batched import BOM, chained, file_mix

pytestmark = pytest.mark.gpu

QUERIES = ["count", "select(.status == 500) | count", ".", "{h: .host, s: .status}", "sum(.bytes)",
           "group_by(.host) | count", "select(.status >= 404) | {h: .host}"]


def write(tmp_path, files, prefix="f"):
    paths = []
    for i, f in enumerate(files):
        p = tmp_path / f"{prefix}{i:04d}.ndjson"
        p.write_bytes(f)
        paths.append(str(p))
    return paths


class Fifo:
    """A named pipe fed by a writer thread (a non-seekable input)."""

    def __init__(self, tmp_path, data, name="pipe"):
        self.path = str(tmp_path / name)
        os.mkfifo(self.path)
        self.t = threading.Thread(target=self._feed, args=(data,), daemon=True)
        self.t.start()

    def _feed(self, data):
        with open(self.path, "wb") as f:
            for i in range(0, len(data), 1 << 16):  # short writes, like a producer
                f.write(data[i:i + (1 << 16)])

    def join(self):
        self.t.join(timeout=60)


def test_paths_match_the_stream(tmp_path):
    files = file_mix(11, 80, 40)
    paths = write(tmp_path, files)
    stream = chained(files)
    for q in QUERIES:
        for fmt in (w.FMT_NDJSON, w.FMT_CSV, w.FMT_COUNT):
            try:
                cpu, cst = oracle.run(q, stream, fmt=fmt, chunk_bytes=1 << 16, threads=4)
            except oracle.RunError:
                continue
            gpu, gst = w.run_paths(q, paths, fmt=fmt, chunk_bytes=1 << 16)
            assert gpu == cpu, (q, fmt)
            for k in ("lines_in", "malformed", "type_errors"):
                assert gst[k] == cst[k], (q, fmt, k)


def test_large_file_cut_into_chunks(tmp_path):
    data = corpus.nginx(5, 20000)
    paths = write(tmp_path, [data, BOM + data[:5000], b"", data[:777].rstrip(b"\n")])
    stream = chained([data, BOM + data[:5000], b"", data[:777].rstrip(b"\n")])
    for q in ("select(.status == 500) | count", "select(.status >= 500) | {p: .path, b: .bytes}",
              "group_by(.host) | sum(.bytes)"):
        cpu, _ = oracle.run(q, stream, threads=4)
        gpu, st = w.run_paths(q, paths, chunk_bytes=1 << 18)
        assert gpu == cpu, q
        assert st["chunks"] >= 4


def test_pipe_streamed_with_long_lines(tmp_path):
    # lines far longer than the chunk arrive through a pipe: each runs alone
    # (gpu/mod.rs:411-458); the rest is carried between reads
    big = b'{"status":500,"bytes":7,"ua":"' + b"x" * 300000 + b'"}\n'
    data = BOM + corpus.nginx(9, 3000) + big + corpus.nginx(10, 50) + big + b'{"status":500}'
    stream = chained([data])
    for i, q in enumerate(("select(.status == 500) | count", "select(.status == 500) | {b: .bytes}", "sum(.bytes)")):
        fifo = Fifo(tmp_path, data, name="p%d" % i)
        gpu, st = w.run_paths(q, [fifo.path], chunk_bytes=1 << 16)
        fifo.join()
        cpu, _ = oracle.run(q, stream, streamed=True)
        assert gpu == cpu, q
        assert st["bytes_in"] == len(data) - 3


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(w.RunError) as e:
        w.run_paths("count", [str(tmp_path / "nope.ndjson")])
    assert e.value.code == 3 and "could not open" in str(e.value)


def cli(*args, data=None):
    return subprocess.run([w.CLI_PATH, *args], input=data, capture_output=True, timeout=120)


def test_cli_stdin_files_and_exit_codes(tmp_path):
    data = corpus.hostile(21, 3000, 0.02)
    cut = data.rindex(b"\n", 0, 40000) + 1  # files split at a line end read as `data`
    paths = write(tmp_path, [data[:cut], data[cut:]])
    for q, fmt, flag in ((".", oracle.FMT_NDJSON, []), ("{h: .host, s: .status}", oracle.FMT_CSV, ["--csv"]),
                         ("select(.status == 500)", oracle.FMT_COUNT, ["-c"]), ("sum(.bytes)", oracle.FMT_NDJSON, [])):
        cpu, _ = oracle.run(q, data, fmt=fmt, threads=4)
        r = cli(q, *flag, data=data)  # stdin
        assert r.returncode == 0, r.stderr
        assert r.stdout == cpu, q
        r2 = cli(q, *flag, *paths)
        assert r2.returncode == 0 and r2.stdout == cpu, q
        r3 = cli(q, *flag, paths[0], "-", data=data[cut:])  # a file, then stdin
        assert r3.returncode == 0 and r3.stdout == cpu, q
    # malformed lines are skipped with a summary on stderr (run.rs:104-121)
    r = cli("count", data=data)
    assert b"malformed line" in r.stderr
    assert cli("count", "--skip-invalid", data=data).stderr == b""
    # no rows from a non-aggregate query: exit 1, like grep (run.rs:95-101)
    r = cli("select(.status == 999)", data=data)
    assert r.returncode == 1 and r.stdout == b""
    assert cli("select(.status == 999) | count", data=data).returncode == 0
    # --stats (run.rs:122-150)
    r = cli("count", "--stats", "--skip-invalid", *paths)
    assert r.returncode == 0 and b"GB/s" in r.stderr and b"lines read" in r.stderr


def test_cli_strict_names_the_line(tmp_path):
    data = b'{"a":1}\n{"a":2}\n\n{"a":}\n{"a":5}\n'
    p = write(tmp_path, [data])[0]
    r = cli("count", "--strict", p)
    assert r.returncode == 1
    assert b"malformed JSON on line 4 of " + p.encode() in r.stderr
    r = cli(".a.b", "--strict", data=b'{"a":{"b":1}}\n{"a":1}\n')  # a type error stops too (exec/cpu.rs:401-405)
    assert r.returncode == 1 and b"line 2" in r.stderr
    r = cli("count", "/nonexistent/x.ndjson")
    assert r.returncode == 1 and b"could not open" in r.stderr
