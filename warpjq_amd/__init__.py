"""warpjq's GPU backend for AMD Instinct MI355X (gfx950).

Python access to libwarpjq_amd.so: the reference's device ABI
(include/warpjq_kernels.h, the ten entry points warpjq's Rust host binds in
crates/warpjq-core/src/gpu/ffi.rs:149-184) plus the additive exports of
include/warpjq_amd.h (the chunked host driver, HBM-resident submit, exact
aggregate parts, device timing).

Everything here calls the HIP library; there is no CPU path. Loading fails
loudly when the library has not been built.
"""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwarpjq_amd.so")

FMT_NDJSON, FMT_CSV, FMT_COUNT = 0, 1, 2
WARN, SKIP, ABORT = 0, 1, 2
OUT_AGG = 3  # warpjq_out_kind (warpjq_kernels.h)
GPU_CHUNK = 64 << 20          # gpu/mod.rs:57
MAX_LINE = 64 << 20           # chunk.rs:22
NONE = 0xFFFFFFFF

STATUS_NAMES = {0: "ok", 1: "hip error", 2: "no device", 3: "out of memory", 4: "abi mismatch",
                5: "invalid argument"}

u32, u64, f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double
P = ctypes.POINTER


class Step(ctypes.Structure):
    _fields_ = [("kind", u32), ("index", u32), ("key_off", u32), ("key_len", u32), ("key_hash", u32)]


class PathT(ctypes.Structure):
    _fields_ = [("step_off", u32), ("step_count", u32)]


class Cmp(ctypes.Structure):
    _fields_ = [("path", u32), ("op", u32), ("lit_kind", u32), ("lit_off", u32), ("lit_len", u32),
                ("_pad", u32), ("lit_num", f64)]


class CondOp(ctypes.Structure):
    _fields_ = [("op", u32), ("arg", u32)]


class Program(ctypes.Structure):
    _fields_ = [("steps", P(Step)), ("n_steps", u32), ("paths", P(PathT)), ("n_paths", u32),
                ("cmps", P(Cmp)), ("n_cmps", u32), ("cond_rpn", P(CondOp)), ("n_cond", u32),
                ("blob", P(ctypes.c_uint8)), ("n_blob", u32), ("needed_paths", P(u32)), ("n_needed", u32),
                ("slot_of_path", P(u32)), ("output_kind", u32), ("output_path", u32), ("agg_kind", u32),
                ("agg_path", u32), ("group_path", u32), ("has_filter", u32), ("n_fields", u32),
                ("field_paths", P(u32)), ("prefix_off", P(u32)), ("prefix_len", P(u32)),
                ("suffix_off", u32), ("suffix_len", u32), ("csv_mode", u32)]


class AggPartial(ctypes.Structure):
    _fields_ = [("count", u64), ("numeric", u64), ("sum", f64), ("min", f64), ("max", f64),
                ("saw_non_numeric", u32), ("_pad", u32)]


class Group(ctypes.Structure):
    _fields_ = [("key_off", u32), ("key_len", u32), ("key_kind", u32), ("_pad", u32), ("count", u64),
                ("numeric", u64), ("sum", f64), ("min", f64), ("max", f64)]


class ChunkResult(ctypes.Structure):
    _fields_ = [("n_lines", u64), ("n_blank", u64), ("n_invalid", u64), ("n_type_error", u64),
                ("n_fallback", u64), ("n_selected", u64), ("out_bytes", P(ctypes.c_uint8)), ("out_len", u64),
                ("out_line_idx", P(u32)), ("out_row_off", P(u64)), ("fallback_idx", P(u32)),
                ("fallback_off", P(u32)), ("fallback_len", P(u32)), ("chunk_overflow", u32),
                ("first_invalid", u32), ("agg", AggPartial), ("groups", P(Group)), ("n_groups", u32),
                ("group_overflow", u32)]


class AggExt(ctypes.Structure):
    _fields_ = [("isum_lo", u64), ("isum_hi", u64), ("fsum", f64), ("fcomp", f64)]


class Span(ctypes.Structure):
    _fields_ = [("d_bytes", ctypes.c_void_p), ("n_bytes", u64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("budget", u64), ("reserved", u64), ("loads", u64), ("evictions", u64), ("completed", u64),
                ("resident", u64), ("p50_ms", f64), ("p99_ms", f64), ("max_ms", f64), ("engines_opened", u64),
                ("engines_reprogrammed", u64)]


class Stats(ctypes.Structure):
    _fields_ = [("bytes_in", u64), ("lines_in", u64), ("lines_out", u64), ("malformed", u64),
                ("type_errors", u64), ("gpu_fallback_lines", u64), ("gpu_redone_chunks", u64),
                ("h2d_ms", f64), ("kernel_ms", f64), ("chunks", u64), ("err_line", u64)]


# The exported C symbols of include/warpjq_kernels.h and include/warpjq_amd.h.
REFERENCE_SYMBOLS = ["warpjq_probe", "warpjq_abi_check", "warpjq_ctx_create", "warpjq_ctx_destroy",
                     "warpjq_slot_buffer", "warpjq_slot_capacity", "warpjq_max_lines", "warpjq_submit",
                     "warpjq_wait", "warpjq_device_name"]
ADDITIVE_SYMBOLS = ["warpjq_amd_ext_value", "warpjq_amd_wait_ext", "warpjq_amd_submit_device",
                    "warpjq_amd_slot_timing", "warpjq_amd_kernel_width", "warpjq_amd_set_claim", "warpjq_amd_set_arena",
                    "warpjq_amd_grid", "warpjq_amd_submit_spans", "warpjq_amd_fallback_spans",
                    "warpjq_amd_group_keys", "warpjq_amd_submit_pieces", "warpjq_amd_run_buffer", "warpjq_amd_run_csv",
                    "warpjq_amd_run_files", "warpjq_amd_device_upload", "warpjq_amd_device_free",
                    "warpjq_amd_server_create", "warpjq_amd_server_destroy", "warpjq_amd_server_add_dataset",
                    "warpjq_amd_server_submit", "warpjq_amd_server_wait", "warpjq_amd_server_get_stats",
                    "warpjq_amd_ctx_reprogram", "warpjq_amd_set_graphs", "warpjq_amd_autotune",
                    "warpjq_amd_free", "warpjq_amd_compile", "warpjq_amd_program_view",
                    "warpjq_amd_program_free", "warpjq_amd_gen_nginx", "warpjq_amd_gen_nested",
                    "warpjq_amd_first_type_error", "warpjq_amd_noalloc_max_bytes",
                    "warpjq_amd_aggregate_device_noalloc", "warpjq_amd_set_group_table",
                    "warpjq_amd_group_table", "warpjq_amd_run_paths"]


def build(jobs=8):
    """Compiles libwarpjq_amd.so for gfx950 in-tree (hipcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", os.path.join(_HERE, "csrc"), f"-j{jobs}"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run warpjq_amd.build() (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        cp, sz = ctypes.c_char_p, ctypes.c_size_t
        L.warpjq_probe.argtypes = [cp, sz]
        L.warpjq_abi_check.argtypes = [P(u64), sz]
        L.warpjq_ctx_create.argtypes = [P(Program), u64, u32, P(ctypes.c_void_p), cp, sz]
        L.warpjq_ctx_destroy.argtypes = [ctypes.c_void_p]
        L.warpjq_ctx_destroy.restype = None
        L.warpjq_slot_buffer.argtypes = [ctypes.c_void_p, u32]
        L.warpjq_slot_buffer.restype = ctypes.c_void_p
        L.warpjq_slot_capacity.argtypes = [ctypes.c_void_p]
        L.warpjq_slot_capacity.restype = u64
        L.warpjq_max_lines.argtypes = [ctypes.c_void_p]
        L.warpjq_max_lines.restype = u64
        L.warpjq_submit.argtypes = [ctypes.c_void_p, u32, u64, cp, sz]
        L.warpjq_wait.argtypes = [ctypes.c_void_p, u32, P(ChunkResult), cp, sz]
        L.warpjq_device_name.argtypes = [cp, sz]
        L.warpjq_amd_ext_value.argtypes = [P(AggExt)]
        L.warpjq_amd_ext_value.restype = f64
        L.warpjq_amd_wait_ext.argtypes = [ctypes.c_void_p, u32, P(AggExt), P(P(AggExt))]
        L.warpjq_amd_submit_device.argtypes = [ctypes.c_void_p, u32, ctypes.c_void_p, u64, cp, sz]
        L.warpjq_amd_slot_timing.argtypes = [ctypes.c_void_p, u32, P(ctypes.c_float), P(ctypes.c_float),
                                             P(ctypes.c_float)]
        L.warpjq_amd_kernel_width.argtypes = [ctypes.c_void_p]
        L.warpjq_amd_kernel_width.restype = u32
        L.warpjq_amd_set_claim.argtypes = [ctypes.c_void_p, u32]
        L.warpjq_amd_set_arena.argtypes = [ctypes.c_void_p, u32]
        L.warpjq_amd_grid.argtypes = [ctypes.c_void_p]
        L.warpjq_amd_submit_spans.argtypes = [ctypes.c_void_p, u32, P(Span), u32, cp, sz]
        L.warpjq_amd_fallback_spans.argtypes = [ctypes.c_void_p, u32]
        L.warpjq_amd_fallback_spans.restype = P(u32)
        L.warpjq_amd_group_keys.argtypes = [ctypes.c_void_p, u32]
        L.warpjq_amd_group_keys.restype = P(ctypes.c_uint8)
        L.warpjq_amd_grid.restype = u32
        L.warpjq_amd_run_buffer.argtypes = [cp, ctypes.c_void_p, u64, ctypes.c_int, ctypes.c_int, u64, u64,
                                            P(P(ctypes.c_uint8)), P(u64), P(Stats), cp, sz]
        L.warpjq_amd_run_buffer.restype = ctypes.c_int
        L.warpjq_amd_run_csv.argtypes = L.warpjq_amd_run_buffer.argtypes
        L.warpjq_amd_run_csv.restype = ctypes.c_int
        L.warpjq_amd_run_files.argtypes = [cp, P(ctypes.c_void_p), P(u64), u32, ctypes.c_int, ctypes.c_int, u64,
                                           u64, P(P(ctypes.c_uint8)), P(u64), P(Stats), cp, sz]
        L.warpjq_amd_run_files.restype = ctypes.c_int
        L.warpjq_amd_run_paths.argtypes = [cp, P(cp), u32, ctypes.c_int, ctypes.c_int, u64, u64,
                                           P(P(ctypes.c_uint8)), P(u64), P(Stats), cp, sz]
        L.warpjq_amd_run_paths.restype = ctypes.c_int
        L.warpjq_amd_submit_pieces.argtypes = [ctypes.c_void_p, u32, u64, ctypes.c_void_p, u32, cp, sz]
        L.warpjq_amd_ctx_reprogram.argtypes = [ctypes.c_void_p, P(Program), cp, sz]
        L.warpjq_amd_set_graphs.argtypes = [ctypes.c_void_p, u32]
        L.warpjq_amd_autotune.argtypes = [ctypes.c_void_p, ctypes.c_void_p, u64, P(u32), P(u32),
                                          P(ctypes.c_float), cp, sz]
        L.warpjq_amd_device_upload.argtypes = [ctypes.c_void_p, u64, cp, sz]
        L.warpjq_amd_device_upload.restype = ctypes.c_void_p
        L.warpjq_amd_noalloc_max_bytes.argtypes = []
        L.warpjq_amd_noalloc_max_bytes.restype = u64
        L.warpjq_amd_aggregate_device_noalloc.argtypes = [P(Program), ctypes.c_void_p, u64, P(ChunkResult),
                                                          P(AggExt), cp, sz]
        L.warpjq_amd_aggregate_device_noalloc.restype = ctypes.c_int
        L.warpjq_amd_device_free.argtypes = [ctypes.c_void_p]
        L.warpjq_amd_device_free.restype = None
        L.warpjq_amd_server_create.argtypes = [u64, u32, u64, cp, sz]
        L.warpjq_amd_server_create.restype = ctypes.c_void_p
        L.warpjq_amd_server_destroy.argtypes = [ctypes.c_void_p]
        L.warpjq_amd_server_destroy.restype = None
        L.warpjq_amd_server_add_dataset.argtypes = [ctypes.c_void_p, ctypes.c_void_p, u64]
        L.warpjq_amd_server_add_dataset.restype = ctypes.c_int64
        L.warpjq_amd_server_submit.argtypes = [ctypes.c_void_p, cp, P(u32), u32, ctypes.c_int, ctypes.c_int]
        L.warpjq_amd_server_submit.restype = ctypes.c_int64
        L.warpjq_amd_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_int64, P(P(ctypes.c_uint8)), P(u64),
                                             P(Stats), P(f64), cp, sz]
        L.warpjq_amd_server_wait.restype = ctypes.c_int
        L.warpjq_amd_server_get_stats.argtypes = [ctypes.c_void_p, P(ServerStats)]
        L.warpjq_amd_server_get_stats.restype = None
        L.warpjq_amd_free.argtypes = [ctypes.c_void_p]
        L.warpjq_amd_free.restype = None
        L.warpjq_amd_compile.argtypes = [cp, ctypes.c_int, cp, sz]
        L.warpjq_amd_compile.restype = ctypes.c_void_p
        L.warpjq_amd_program_view.argtypes = [ctypes.c_void_p]
        L.warpjq_amd_program_view.restype = P(Program)
        L.warpjq_amd_program_free.argtypes = [ctypes.c_void_p]
        L.warpjq_amd_program_free.restype = None
        L.warpjq_amd_gen_nginx.argtypes = [u64, u64, u32, u64, P(P(ctypes.c_uint8)), P(u64), P(u64), P(u64)]
        L.warpjq_amd_gen_nginx.restype = ctypes.c_int
        L.warpjq_amd_gen_nested.argtypes = [u64, u64, u32, u64, P(P(ctypes.c_uint8)), P(u64), P(u64), P(u64)]
        L.warpjq_amd_gen_nested.restype = ctypes.c_int
        _lib = L
    return _lib


class DeviceError(RuntimeError):
    pass


class RunError(RuntimeError):
    def __init__(self, code, line, msg):
        super().__init__(msg)
        self.code = code
        self.line = line


def _err():
    return ctypes.create_string_buffer(512)


def probe():
    e = _err()
    rc = lib().warpjq_probe(e, 512)
    if rc != 0:
        raise DeviceError(f"{STATUS_NAMES.get(rc, rc)}: {e.value.decode()}")


def device_name():
    b = ctypes.create_string_buffer(256)
    if lib().warpjq_device_name(b, 256) != 0:
        return None
    return b.value.decode()


def abi_sizes():
    return [ctypes.sizeof(t) for t in (Step, PathT, Cmp, CondOp, AggPartial, Group)]


def abi_check(sizes=None):
    s = (u64 * 6)(*(sizes or abi_sizes()))
    return lib().warpjq_abi_check(s, 6)


def _addr(data, n=None):
    """(address, length) of bytes, bytearray or a numpy uint8 array."""
    if isinstance(data, bytes):
        return ctypes.cast(ctypes.c_char_p(data), ctypes.c_void_p).value, len(data) if n is None else n
    if hasattr(data, "ctypes"):
        return data.ctypes.data, data.nbytes if n is None else n
    buf = (ctypes.c_char * len(data)).from_buffer(data)
    return ctypes.addressof(buf), len(data) if n is None else n


def run_files(query, files, fmt=FMT_NDJSON, on_invalid=SKIP, chunk_bytes=GPU_CHUNK, max_line=MAX_LINE):
    """Runs `query` over a list of inputs read as one stream (the reference's
    multi-file form) on the GPU. Returns (output bytes, stats)."""
    L = lib()
    keep = list(files)
    n = len(keep)
    ptrs = (ctypes.c_void_p * max(1, n))()
    lens = (u64 * max(1, n))()
    for i, f in enumerate(keep):
        ptrs[i], lens[i] = _addr(f)
    out = P(ctypes.c_uint8)()
    out_len = u64()
    st = Stats()
    e = ctypes.create_string_buffer(2048)
    q = query.encode() if isinstance(query, str) else query
    rc = L.warpjq_amd_run_files(q, ptrs, lens, n, fmt, on_invalid, chunk_bytes, max_line, ctypes.byref(out),
                                ctypes.byref(out_len), ctypes.byref(st), e, 2048)
    del keep
    return _finish(L, rc, out, out_len, st, e)


def run_paths(query, paths, fmt=FMT_NDJSON, on_invalid=SKIP, chunk_bytes=GPU_CHUNK, max_line=MAX_LINE):
    """Runs `query` over files named by path, read as one stream ("-" or an
    empty list reads stdin), on the GPU. Returns (output bytes, stats)."""
    L = lib()
    enc = [p.encode() if isinstance(p, str) else p for p in paths]
    arr = (ctypes.c_char_p * max(1, len(enc)))(*enc)
    out = P(ctypes.c_uint8)()
    out_len = u64()
    st = Stats()
    e = ctypes.create_string_buffer(2048)
    q = query.encode() if isinstance(query, str) else query
    rc = L.warpjq_amd_run_paths(q, arr, len(enc), fmt, on_invalid, chunk_bytes, max_line, ctypes.byref(out),
                                ctypes.byref(out_len), ctypes.byref(st), e, 2048)
    return _finish(L, rc, out, out_len, st, e)


CLI_PATH = os.path.join(_HERE, "warpjq-amd")


def run(query, data, fmt=FMT_NDJSON, on_invalid=SKIP, chunk_bytes=GPU_CHUNK, max_line=MAX_LINE, n=None):
    """Runs `query` over `data` (bytes, or a numpy uint8 array whose first
    `n` bytes are the input) on the GPU. Returns (output bytes, stats)."""
    L = lib()
    keep = data
    addr, nbytes = _addr(data, n)
    out = P(ctypes.c_uint8)()
    out_len = u64()
    st = Stats()
    e = ctypes.create_string_buffer(2048)
    q = query.encode() if isinstance(query, str) else query
    rc = L.warpjq_amd_run_buffer(q, addr, nbytes, fmt, on_invalid, chunk_bytes, max_line, ctypes.byref(out),
                                 ctypes.byref(out_len), ctypes.byref(st), e, 2048)
    del keep
    return _finish(L, rc, out, out_len, st, e)


def run_csv(query, data, fmt=FMT_NDJSON, on_invalid=SKIP, chunk_bytes=GPU_CHUNK, max_line=MAX_LINE):
    """Runs `query` over a CSV input (header line, then one record per line,
    transcoded to NDJSON objects on the device; warpjq_amd_run_csv).
    Returns (output bytes, stats)."""
    L = lib()
    keep = data
    addr, nbytes = _addr(data)
    out = P(ctypes.c_uint8)()
    out_len = u64()
    st = Stats()
    e = ctypes.create_string_buffer(2048)
    q = query.encode() if isinstance(query, str) else query
    rc = L.warpjq_amd_run_csv(q, addr, nbytes, fmt, on_invalid, chunk_bytes, max_line, ctypes.byref(out),
                              ctypes.byref(out_len), ctypes.byref(st), e, 2048)
    del keep
    return _finish(L, rc, out, out_len, st, e)


def _finish(L, rc, out, out_len, st, e):
    stats = {k: getattr(st, k) for k, _ in Stats._fields_}
    if rc != 0:
        raise RunError(rc, st.err_line, e.value.decode(errors="replace"))
    b = ctypes.string_at(out, out_len.value) if out_len.value else b""
    L.warpjq_amd_free(out)
    return b, stats


class Compiled:
    """A lowered program (warpjq_amd_compile)."""

    def __init__(self, query, csv=False):
        e = ctypes.create_string_buffer(2048)
        self.h = lib().warpjq_amd_compile(query.encode(), 1 if csv else 0, e, 2048)
        if not self.h:
            raise ValueError(e.value.decode())
        self.view = lib().warpjq_amd_program_view(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().warpjq_amd_program_free(self.h)
            self.h = None


class Context:
    """A device context over the raw ABI (warpjq_ctx_create ... destroy)."""

    def __init__(self, query, chunk_bytes=GPU_CHUNK, n_slots=2, csv=False):
        self.prog = Compiled(query, csv)
        self.ctx = ctypes.c_void_p()
        e = _err()
        rc = lib().warpjq_ctx_create(self.prog.view, chunk_bytes, n_slots, ctypes.byref(self.ctx), e, 512)
        if rc != 0:
            raise DeviceError(f"ctx_create: {STATUS_NAMES.get(rc, rc)}: {e.value.decode()}")
        self.n_slots = n_slots
        self._claim, self._arena = 8 << 10, 18 << 10  # the library's defaults

    def close(self):
        if self.ctx:
            lib().warpjq_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def reprogram(self, query, csv=False):
        """Points this (idle) context at another query with the same output shape."""
        prog = Compiled(query, csv)
        e = _err()
        rc = lib().warpjq_amd_ctx_reprogram(self.ctx, prog.view, e, 512)
        if rc != 0:
            raise DeviceError(f"reprogram: {STATUS_NAMES.get(rc, rc)}: {e.value.decode()}")
        self.prog = prog

    def autotune(self, dptr, n, cache=None):
        """Times claim / arena pairs on resident bytes and keeps the fastest.
        With `cache` (a TuneCache), a stored choice for the same device,
        kernel shape and line-length class is applied instead of sweeping.
        Returns (claim, arena, main_ms or None when taken from the cache)."""
        key = None
        if cache is not None:
            key = cache.key(self, dptr, n)
            hit = cache.get(key)
            if hit:
                # the arena first when it grows, so the claim always fits
                if hit[1] >= self._arena:
                    self.set_arena(hit[1])
                    self.set_claim(hit[0])
                else:
                    self.set_claim(hit[0])
                    self.set_arena(hit[1])
                self._claim, self._arena = hit[0], hit[1]
                if len(hit) > 2 and hit[2]:
                    self.set_group_table(hit[2])
                return hit[0], hit[1], None
        c, a, ms = u32(), u32(), ctypes.c_float()
        e = _err()
        rc = lib().warpjq_amd_autotune(self.ctx, dptr, n, ctypes.byref(c), ctypes.byref(a), ctypes.byref(ms), e, 512)
        if rc != 0:
            raise DeviceError(f"autotune: {STATUS_NAMES.get(rc, rc)}: {e.value.decode()}")
        self._claim, self._arena = c.value, a.value
        if cache is not None:
            pv = self.prog.view.contents
            grouped = pv.output_kind == OUT_AGG and pv.group_path != 0xFFFFFFFF
            cache.put(key, (c.value, a.value, self.group_table if grouped else 0))
        return c.value, a.value, ms.value

    @property
    def group_table(self):
        """Group-table slots in use (grouped contexts)."""
        L = lib()
        L.warpjq_amd_group_table.restype = u32
        return L.warpjq_amd_group_table(self.ctx)

    def set_group_table(self, slots):
        if lib().warpjq_amd_set_group_table(self.ctx, slots) != 0:
            raise ValueError("group table: a power of two from 4096 up to the context's allocation, grouped queries")

    def set_graphs(self, per_slot):
        """Cache up to `per_slot` captured chains (hipGraph) per slot for resident submits."""
        if lib().warpjq_amd_set_graphs(self.ctx, per_slot) != 0:
            raise ValueError("set_graphs: at most 65536 per slot")

    def set_claim(self, nbytes):
        if lib().warpjq_amd_set_claim(self.ctx, nbytes) != 0:
            raise ValueError("claim must be 1-8 KiB in 1 KiB steps and fit the arena with 1 KiB to spare")
        self._claim = nbytes

    def set_arena(self, nbytes):
        if lib().warpjq_amd_set_arena(self.ctx, nbytes) != 0:
            raise ValueError("arena must be a multiple of 256 B, at most 32 KiB, at least the claim plus 1 KiB")
        self._arena = nbytes

    @property
    def grid(self):
        return lib().warpjq_amd_grid(self.ctx)

    @property
    def capacity(self):
        return lib().warpjq_slot_capacity(self.ctx)

    @property
    def max_lines(self):
        return lib().warpjq_max_lines(self.ctx)

    def fill(self, slot, data):
        buf = lib().warpjq_slot_buffer(self.ctx, slot)
        ctypes.memmove(buf, data, len(data))

    def submit(self, slot, n):
        e = _err()
        rc = lib().warpjq_submit(self.ctx, slot, n, e, 512)
        if rc != 0:
            raise DeviceError(f"submit: {STATUS_NAMES.get(rc, rc)}: {e.value.decode()}")

    def submit_device(self, slot, dptr, n):
        e = _err()
        rc = lib().warpjq_amd_submit_device(self.ctx, slot, dptr, n, e, 512)
        if rc != 0:
            raise DeviceError(f"submit_device: {STATUS_NAMES.get(rc, rc)}: {e.value.decode()}")

    def submit_spans(self, slot, spans):
        """One launch over device spans [(dptr, n_bytes), ...] (the batched form)."""
        arr = (Span * max(1, len(spans)))()
        for i, (p, n) in enumerate(spans):
            arr[i].d_bytes = p
            arr[i].n_bytes = n
        e = _err()
        rc = lib().warpjq_amd_submit_spans(self.ctx, slot, arr, len(spans), e, 512)
        if rc != 0:
            raise DeviceError(f"submit_spans: {STATUS_NAMES.get(rc, rc)}: {e.value.decode()}")

    def fallback_spans(self, slot, n):
        """Span index of each of the `n` fallback lines of the last wait on `slot`."""
        p = lib().warpjq_amd_fallback_spans(self.ctx, slot)
        return [p[i] for i in range(n)]

    def group_key(self, slot, off, n):
        """Key bytes of a group of a multi-span submit (key_off indexes the key blob)."""
        return ctypes.string_at(ctypes.addressof(lib().warpjq_amd_group_keys(self.ctx, slot).contents) + off, n)

    def wait(self, slot):
        r = ChunkResult()
        e = _err()
        rc = lib().warpjq_wait(self.ctx, slot, ctypes.byref(r), e, 512)
        if rc != 0:
            raise DeviceError(f"wait: {STATUS_NAMES.get(rc, rc)}: {e.value.decode()}")
        return r

    def first_type_error(self, slot):
        """Chunk-relative index of the first type-error line of the last wait
        on `slot` (None when there was none)."""
        L = lib()
        L.warpjq_amd_first_type_error.restype = ctypes.c_uint32
        v = L.warpjq_amd_first_type_error(self.ctx, slot)
        return None if v == 0xFFFFFFFF else v

    def ext(self, slot):
        x = AggExt()
        g = P(AggExt)()
        lib().warpjq_amd_wait_ext(self.ctx, slot, ctypes.byref(x), ctypes.byref(g))
        return x, g

    def timing(self, slot):
        """(h2d_ms, chain_ms, main_kernel_ms) of the last submit on `slot`."""
        a, b, m = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        lib().warpjq_amd_slot_timing(self.ctx, slot, ctypes.byref(a), ctypes.byref(b), ctypes.byref(m))
        return a.value, b.value, m.value


class TuneCache:
    """Autotuner choices persisted as JSON: (device, kernel width, output
    shape, line-length class) -> (claim, arena, group-table slots or 0)."""

    def __init__(self, path):
        import json
        self.path = path
        try:
            with open(path) as f:
                self.d = json.load(f)
        except (OSError, ValueError):
            self.d = {}

    def key(self, ctx, dptr, n):
        # line-length class from one pass over (a prefix of) the bytes
        m = min(n, 16 << 20)
        ctx.submit_device(0, dptr, m)
        r = ctx.wait(0)
        avg = m / max(1, r.n_lines)
        cls = 0
        while (64 << cls) < avg and cls < 16:
            cls += 1
        pv = ctx.prog.view.contents
        shape = "rows" if pv.output_kind != OUT_AGG else ("groups" if pv.group_path != 0xFFFFFFFF else "agg")
        return f"{device_name()}|w{lib().warpjq_amd_kernel_width(ctx.ctx)}|{shape}|{64 << cls}"

    def get(self, key):
        v = self.d.get(key)
        return tuple(v) if v else None

    def put(self, key, v):
        import json
        self.d[key] = list(v)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.d, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def aggregate_device_noalloc(query, dptr, n):
    """An ungrouped aggregate over n device bytes through
    warpjq_amd_aggregate_device_noalloc (no device allocation per call).
    Returns (ChunkResult, AggExt)."""
    prog = Compiled(query)
    r = ChunkResult()
    x = AggExt()
    e = _err()
    rc = lib().warpjq_amd_aggregate_device_noalloc(prog.view, dptr, n, ctypes.byref(r), ctypes.byref(x), e, 512)
    if rc != 0:
        raise DeviceError(f"aggregate_device_noalloc: {STATUS_NAMES.get(rc, rc)}: {e.value.decode()}")
    return r, x


class DeviceBuffer:
    """Host bytes copied to a new device allocation (+256 newline bytes)."""

    def __init__(self, data):
        addr, n = _addr(data)
        e = _err()
        self.ptr = lib().warpjq_amd_device_upload(addr, n, e, 512)
        if not self.ptr:
            raise DeviceError(f"device_upload: {e.value.decode()}")
        self.n = n

    def free(self):
        if self.ptr:
            lib().warpjq_amd_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


class Server:
    """The served form (warpjq_amd_server_*): resident datasets under an HBM
    budget, requests run by `workers` threads as they are admitted."""

    def __init__(self, hbm_budget, workers=4, chunk_bytes=GPU_CHUNK):
        e = _err()
        self.h = lib().warpjq_amd_server_create(hbm_budget, workers, chunk_bytes, e, 512)
        if not self.h:
            raise DeviceError(e.value.decode())
        self._keep = []

    def add_dataset(self, data):
        """Registers bytes (kept referenced for the server's life)."""
        self._keep.append(data)
        addr, n = _addr(data)
        i = lib().warpjq_amd_server_add_dataset(self.h, addr, n)
        if i < 0:
            raise ValueError("add_dataset failed")
        return i

    def submit(self, query, datasets, fmt=FMT_NDJSON, on_invalid=SKIP):
        ids = (u32 * max(1, len(datasets)))(*datasets)
        q = query.encode() if isinstance(query, str) else query
        r = lib().warpjq_amd_server_submit(self.h, q, ids, len(datasets), fmt, on_invalid)
        if r < 0:
            raise ValueError("server_submit: bad argument")
        return r

    def wait(self, req):
        """(output bytes, stats, latency_ms); raises RunError like run()."""
        L = lib()
        out = P(ctypes.c_uint8)()
        out_len = u64()
        st = Stats()
        lat = f64()
        e = ctypes.create_string_buffer(2048)
        rc = L.warpjq_amd_server_wait(self.h, req, ctypes.byref(out), ctypes.byref(out_len), ctypes.byref(st),
                                      ctypes.byref(lat), e, 2048)
        b, stats = _finish(L, rc, out, out_len, st, e)
        return b, stats, lat.value

    def stats(self):
        s = ServerStats()
        lib().warpjq_amd_server_get_stats(self.h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().warpjq_amd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


STATUS_ORDER = (200, 404, 500, 301, 304, 403, 502, 503)


def gen_nginx(target_bytes, seed=1, threads=None, pad=256):
    """Synthetic nginx-schema NDJSON. Returns (numpy uint8 array of n+pad
    bytes, n_bytes, n_lines, {status: count})."""
    import numpy as np
    threads = threads or min(16, os.cpu_count() or 1)
    out = P(ctypes.c_uint8)()
    n, lines = u64(), u64()
    hist = (u64 * 8)()
    rc = lib().warpjq_amd_gen_nginx(seed, target_bytes, threads, pad, ctypes.byref(out), ctypes.byref(n),
                                    ctypes.byref(lines), hist)
    if rc != 0:
        raise MemoryError("generator allocation failed")
    total = n.value + pad
    arr = np.empty(total, dtype=np.uint8)
    ctypes.memmove(arr.ctypes.data, out, total)
    lib().warpjq_amd_free(out)
    return arr, n.value, lines.value, dict(zip(STATUS_ORDER, list(hist)))


def gen_nested(target_bytes, seed=1, threads=None, pad=256):
    """Synthetic nested-schema NDJSON. Returns (numpy uint8 array of n+pad
    bytes, n_bytes, n_lines, n_malformed)."""
    import numpy as np
    threads = threads or min(16, os.cpu_count() or 1)
    out = P(ctypes.c_uint8)()
    n, lines, bad = u64(), u64(), u64()
    rc = lib().warpjq_amd_gen_nested(seed, target_bytes, threads, pad, ctypes.byref(out), ctypes.byref(n),
                                     ctypes.byref(lines), ctypes.byref(bad))
    if rc != 0:
        raise MemoryError("generator allocation failed")
    total = n.value + pad
    arr = np.empty(total, dtype=np.uint8)
    ctypes.memmove(arr.ctypes.data, out, total)
    lib().warpjq_amd_free(out)
    return arr, n.value, lines.value, bad.value


def result_dict(r):
    """Plain copy of a warpjq_chunk_result's scalar fields and arrays."""
    d = {k: getattr(r, k) for k in ("n_lines", "n_blank", "n_invalid", "n_type_error", "n_fallback",
                                    "n_selected", "out_len", "chunk_overflow", "first_invalid", "n_groups",
                                    "group_overflow")}
    d["agg"] = {k: getattr(r.agg, k) for k, _ in AggPartial._fields_ if k != "_pad"}
    d["out"] = ctypes.string_at(r.out_bytes, r.out_len) if r.out_len else b""
    d["line_idx"] = [r.out_line_idx[i] for i in range(r.n_selected)] if r.n_selected else []
    d["row_off"] = [r.out_row_off[i] for i in range(r.n_selected + 1)] if r.n_selected else []
    d["fallback"] = sorted((r.fallback_idx[i], r.fallback_off[i], r.fallback_len[i]) for i in range(r.n_fallback))
    d["groups"] = [{k: getattr(r.groups[i], k) for k, _ in Group._fields_ if k != "_pad"}
                   for i in range(r.n_groups)]
    return d
