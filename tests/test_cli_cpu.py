"""The command line's argument handling and exit codes, without a GPU
(crates/warpjq-cli/src/main.rs:46-96, :155-190, :217-248; run.rs:16-101).
Query errors are reported before any device work, so these run anywhere
the library loads."""
import os
import subprocess

import pytest

import warpjq_amd as w


def cli(*args, data=None):
    if not os.path.exists(w.CLI_PATH):
        pytest.skip("warpjq-amd is not built")
    return subprocess.run([w.CLI_PATH, *args], input=data, capture_output=True, timeout=60)


def test_help():
    r = cli("--help")
    assert r.returncode == 0
    assert b"usage: warpjq-amd" in r.stdout and b"--max-line-bytes" in r.stdout


@pytest.mark.parametrize("args,needle", [
    ((), b"no query given"),
    ((".a", "--csv", "--count"), b"cannot be used with"),
    ((".a", "--strict", "--skip-invalid"), b"cannot be used with"),
    ((".a", "--backend", "tpu"), b"invalid value 'tpu'"),
    ((".a", "--backend", "cpu"), b"only the GPU engine"),
    ((".a", "--frobnicate"), b"unexpected argument"),
    ((".a", "--chunk-size"), b"needs a value"),
])
def test_usage_errors_exit_2(args, needle):
    r = cli(*args)
    assert r.returncode == 2, r
    assert needle in r.stderr, r.stderr


def test_query_errors_exit_2_with_a_caret():
    # error.rs:37-51: message, the query, a caret under the offending byte
    r = cli("select(.a ==)", "/nonexistent")
    assert r.returncode == 2
    lines = r.stderr.decode().splitlines()
    assert lines[0].startswith("warpjq-amd: invalid query:")
    assert lines[1] == "  select(.a ==)"
    assert lines[2].strip() == "^"


@pytest.mark.parametrize("size,ok", [("banana", False), ("-1GB", False), ("inf", False), ("nan", False),
                                     ("0", False)])
def test_bad_sizes_exit_1(size, ok):
    # main.rs:217-248; a zero chunk is refused (run.rs:38-40)
    r = cli(".a", "--chunk-size", size, "/nonexistent")
    assert r.returncode == 1
    assert b"size" in r.stderr or b"greater than zero" in r.stderr


def test_gen_writes_the_seeded_generator(tmp_path):
    out = tmp_path / "g.ndjson"
    r = cli("gen", "--preset", "nginx", "--size", "64KB", "-o", str(out), "--seed", "5")
    assert r.returncode == 0, r.stderr
    data = out.read_bytes()
    host, n, lines, _ = w.gen_nginx(64 << 10, seed=5, threads=16, pad=0)
    assert data == host[:n].tobytes()
    assert data.count(b"\n") == lines
    assert b"wrote" in r.stderr
    r = cli("gen", "--preset", "nope")
    assert r.returncode == 2 and b"unknown preset" in r.stderr
    assert cli("gen", "--list").returncode == 0
