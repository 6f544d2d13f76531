"""ctypes access to the test-only host build of the device line evaluator."""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "libhostsim.so")


class HsLine(ctypes.Structure):
    _fields_ = [("status", ctypes.c_uint32), ("selected", ctypes.c_uint32), ("end", ctypes.c_uint32),
                ("has_num", ctypes.c_uint32), ("is_int", ctypes.c_uint32), ("agg_kind_seen", ctypes.c_uint32),
                ("num", ctypes.c_double), ("ival", ctypes.c_int64), ("gkind", ctypes.c_uint32),
                ("gesc", ctypes.c_uint32), ("row_len", ctypes.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = ctypes.CDLL(_PATH)
        L.hs_new.restype = ctypes.c_void_p
        L.hs_new.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        L.hs_free.argtypes = [ctypes.c_void_p]
        L.hs_eval.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(HsLine),
                              ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_uint64,
                              ctypes.POINTER(ctypes.c_uint64)]
        L.hs_parse_number.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)]
        L.hs_parse_number.restype = ctypes.c_int
        L.hs_parse_number_exact.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
        L.hs_parse_number_exact.restype = ctypes.c_int
        L.hs_slow_decimal.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
        L.hs_slow_decimal.restype = ctypes.c_double
        L.hs_validate_deep.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32]
        L.hs_validate_deep.restype = ctypes.c_int
        L.hs_flat_vs_walk.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32]
        L.hs_flat_vs_walk.restype = ctypes.c_int
        L.hs_validate.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
        L.hs_validate.restype = ctypes.c_int
        _lib = L
    return _lib


class Evaluator:
    def __init__(self, query, csv=False):
        err = ctypes.create_string_buffer(512)
        self.h = lib().hs_new(query.encode(), 1 if csv else 0, err, 512)
        if not self.h:
            raise ValueError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().hs_free(self.h)

    def flat_vs_walk(self, line, lead=0):
        return lib().hs_flat_vs_walk(self.h, line, len(line), lead)

    def eval(self, line):
        o = HsLine()
        cap = len(line) * 8 + 256
        row = ctypes.create_string_buffer(cap)
        gk = ctypes.create_string_buffer(cap)
        gl = ctypes.c_uint64()
        lib().hs_eval(self.h, line, len(line), ctypes.byref(o), row, cap, gk, cap, ctypes.byref(gl))
        rowb = row.raw[:o.row_len] if o.row_len and o.row_len != 2**64 - 1 else b""
        return o, rowb, gk.raw[:gl.value]


def parse_number(tok):
    v = ctypes.c_double()
    ii = ctypes.c_int()
    iv = ctypes.c_int64()
    ok = lib().hs_parse_number(tok, len(tok), ctypes.byref(v), ctypes.byref(ii), ctypes.byref(iv))
    return (v.value, bool(ii.value), iv.value) if ok else None


def parse_number_exact(tok):
    v = ctypes.c_double()
    ok = lib().hs_parse_number_exact(tok, len(tok), ctypes.byref(v))
    return v.value if ok else None


def slow_decimal(tok):
    return lib().hs_slow_decimal(tok, len(tok))


def validate_deep(line, cap=24576):
    return lib().hs_validate_deep(line, len(line), cap)


def validate(line):
    return lib().hs_validate(line, len(line))
