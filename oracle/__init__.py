"""ORACLE — test infrastructure only.

ctypes access to the CPU restatement of warpjq's CPU backend (oracle.cpp).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
as the checker; the product in warpjq_amd/ never imports it.
"""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

FMT_NDJSON, FMT_CSV, FMT_COUNT = 0, 1, 2
SKIP, WARN, ABORT = 1, 0, 2
DEFAULT_CHUNK = 256 << 20
DEFAULT_MAX_LINE = 64 << 20


class _Result(ctypes.Structure):
    _fields_ = [
        ("out", ctypes.POINTER(ctypes.c_uint8)),
        ("out_len", ctypes.c_uint64),
        ("bytes_in", ctypes.c_uint64),
        ("lines_in", ctypes.c_uint64),
        ("lines_out", ctypes.c_uint64),
        ("malformed", ctypes.c_uint64),
        ("type_errors", ctypes.c_uint64),
        ("err_kind", ctypes.c_int32),
        ("_pad", ctypes.c_int32),
        ("err_line", ctypes.c_uint64),
        ("err", ctypes.c_char * 1024),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u8p = ctypes.c_char_p
        L.oracle_run.argtypes = [ctypes.c_char_p, u8p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                 ctypes.POINTER(_Result)]
        L.oracle_run.restype = ctypes.c_int
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_validate.argtypes = [u8p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_validate.restype = ctypes.c_int
        L.oracle_lookup.argtypes = [u8p, ctypes.c_uint64, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_lookup.restype = ctypes.c_int
        L.oracle_format_number.argtypes = [ctypes.c_double, ctypes.c_char_p, ctypes.c_uint64]
        L.oracle_format_number.restype = ctypes.c_int
        for name in ("oracle_unescape", "oracle_json_string", "oracle_csv_cell"):
            f = getattr(L, name)
            f.argtypes = [u8p, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_uint64]
            f.restype = ctypes.c_int64
        L.oracle_compare.argtypes = [u8p, ctypes.c_uint64, ctypes.c_char_p]
        L.oracle_compare.restype = ctypes.c_int
        L.oracle_chunks.argtypes = [u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64]
        L.oracle_chunks.restype = ctypes.c_int64
        L.oracle_lines.argtypes = [u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                   ctypes.c_uint64]
        L.oracle_lines.restype = ctypes.c_int64
        L.oracle_csv_to_ndjson.argtypes = [u8p, ctypes.c_uint64, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)),
                                            ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_csv_to_ndjson.restype = ctypes.c_int
        _lib = L
    return _lib


class RunError(Exception):
    def __init__(self, kind, line, msg):
        super().__init__(msg)
        self.kind = kind
        self.line = line


def run(query, data, fmt=FMT_NDJSON, on_invalid=SKIP, chunk_bytes=DEFAULT_CHUNK, threads=0,
        streamed=False, max_line=DEFAULT_MAX_LINE, raise_errors=True):
    """Runs `query` over `data` with the CPU backend semantics.

    Returns (output bytes, stats dict). streamed=False mirrors
    exec::run_bytes (a mapped file); streamed=True mirrors cpu::run_bytes
    (a cursor).
    """
    L = lib()
    r = _Result()
    if isinstance(query, str):
        query = query.encode()
    L.oracle_run(query, data, len(data), fmt, on_invalid, chunk_bytes, max_line, threads,
                 1 if streamed else 0, ctypes.byref(r))
    out = ctypes.string_at(r.out, r.out_len) if r.out else b""
    if r.out:
        L.oracle_free(r.out)
    stats = dict(bytes_in=r.bytes_in, lines_in=r.lines_in, lines_out=r.lines_out,
                 malformed=r.malformed, type_errors=r.type_errors)
    if r.err_kind and raise_errors:
        raise RunError(r.err_kind, r.err_line, r.err.decode(errors="replace"))
    return out, stats


def csv_to_ndjson(data):
    """The NDJSON a CSV input reads as (header as a blank line, one object per
    record, malformed records as the invalid line "{"); this build's own
    specification, so parity with the reference is unpinned."""
    L = lib()
    out = ctypes.POINTER(ctypes.c_uint8)()
    n = ctypes.c_uint64()
    if L.oracle_csv_to_ndjson(data, len(data), ctypes.byref(out), ctypes.byref(n)) != 0:
        raise ValueError("CSV input: the header line is missing or badly quoted")
    b = ctypes.string_at(out, n.value)
    L.oracle_free(out)
    return b


def run_csv(query, data, **kw):
    """`run` over a CSV input (see csv_to_ndjson)."""
    return run(query, csv_to_ndjson(data), **kw)


def validate(line):
    return lib().oracle_validate(line, len(line), None) == 1


KIND_NAMES = ["missing", "null", "bool", "num", "str", "arr", "obj"]


def lookup(line, path):
    """('found', kind, raw) / ('type_error',) / ('invalid',)."""
    k = ctypes.c_int()
    off = ctypes.c_uint64()
    ln = ctypes.c_uint64()
    rc = lib().oracle_lookup(line, len(line), path.encode(), ctypes.byref(k), ctypes.byref(off),
                             ctypes.byref(ln))
    if rc == 0:
        raw = b"null" if k.value == 0 else line[off.value:off.value + ln.value]
        return ("found", KIND_NAMES[k.value], raw)
    if rc == 1:
        return ("type_error",)
    if rc == 2:
        return ("invalid",)
    raise ValueError("bad path query")


def format_number(v):
    buf = ctypes.create_string_buffer(64)
    lib().oracle_format_number(v, buf, 64)
    return buf.value.decode()


def _bytes_fn(name, data, cap_mult=8):
    cap = len(data) * cap_mult + 16
    buf = ctypes.create_string_buffer(cap)
    n = getattr(lib(), name)(data, len(data), buf, cap)
    if n < 0:
        return None
    return buf.raw[:n]


def unescape(raw):
    return _bytes_fn("oracle_unescape", raw)


def json_string(b):
    return _bytes_fn("oracle_json_string", b)


def csv_cell(b):
    return _bytes_fn("oracle_csv_cell", b)


def compare(line, cmp_query):
    return lib().oracle_compare(line, len(line), cmp_query.encode())


def chunks(data, target, streamed=False, max_line=DEFAULT_MAX_LINE):
    """List of (len, first_line) per chunk, or None on a too-long line."""
    cap = len(data) + 8
    arr = (ctypes.c_uint64 * (3 * cap))()
    k = lib().oracle_chunks(data, len(data), target, max_line, 1 if streamed else 0, arr, cap)
    if k < 0:
        return None
    return [(arr[3 * i + 1], arr[3 * i + 2]) for i in range(k)]


def lines(data, first_line=1):
    cap = len(data) + 8
    arr = (ctypes.c_uint64 * (3 * cap))()
    k = lib().oracle_lines(data, len(data), first_line, arr, cap)
    return [(arr[3 * i + 2], data[arr[3 * i]:arr[3 * i] + arr[3 * i + 1]]) for i in range(k)]
