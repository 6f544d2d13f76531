"""The no-allocation entry point (warpjq_amd_aggregate_device_noalloc)
against the context form it varies (warpjq_ctx_create +
warpjq_amd_submit_device + warpjq_wait) on the same resident bytes: two
calls in a row, the second smaller, every counter and aggregate equal, the
exact sum parts equal; sizes past its fixed scratch refused."""
import pytest

import corpus
import warpjq_amd as w

pytestmark = pytest.mark.gpu

QUERIES = ["count", "select(.status == 500) | count", "sum(.bytes)", "min(.bytes)", "max(.bytes)",
           "avg(.bytes)", 'select(.host == "web-01") | sum(.bytes)', "select(.a.b) | count"]


def _ctx(q, dptr, n):
    c = w.Context(q, chunk_bytes=64 << 20, n_slots=1)
    try:
        c.submit_device(0, dptr, n)
        r = c.wait(0)
        x, _ = c.ext(0)
        return r, x
    finally:
        c.close()


def _same(a, b):
    ra, xa = a
    rb, xb = b
    for f in ("n_lines", "n_blank", "n_invalid", "n_type_error", "n_fallback", "first_invalid", "chunk_overflow"):
        assert getattr(ra, f) == getattr(rb, f), f
    for f in ("count", "numeric", "min", "max", "saw_non_numeric"):
        assert getattr(ra.agg, f) == getattr(rb.agg, f), f
    assert ra.agg.sum == rb.agg.sum or (ra.agg.sum != ra.agg.sum and rb.agg.sum != rb.agg.sum)
    for f in ("isum_lo", "isum_hi", "fsum", "fcomp"):
        assert getattr(xa, f) == getattr(xb, f), f


def test_noalloc_matches_the_context_form_twice():
    big = corpus.nginx(7, 200000) + corpus.hostile(8, 3000, 0.1)
    small = corpus.nginx(9, 3000)
    bufs = [w.DeviceBuffer(big), w.DeviceBuffer(small)]
    try:
        for q in QUERIES:
            for b in bufs:  # the second call's input is smaller
                got = w.aggregate_device_noalloc(q, b.ptr, b.n)
                want = _ctx(q, b.ptr, b.n)
                _same(got, want)
                assert got[0].n_lines > 0
        # unaligned starts and lengths
        for off in (1, 7, 15):
            got = w.aggregate_device_noalloc("sum(.bytes)", bufs[1].ptr + off, bufs[1].n - off)
            want = _ctx("sum(.bytes)", bufs[1].ptr + off, bufs[1].n - off)
            _same(got, want)
    finally:
        for b in bufs:
            b.free()


def test_noalloc_refuses_what_it_cannot_serve():
    b = w.DeviceBuffer(b'{"a":1}\n')
    try:
        with pytest.raises(w.DeviceError):
            w.aggregate_device_noalloc("count", b.ptr, w.lib().warpjq_amd_noalloc_max_bytes() + 1)
        with pytest.raises(w.DeviceError):
            w.aggregate_device_noalloc("group_by(.a) | count", b.ptr, b.n)
        with pytest.raises(w.DeviceError):
            w.aggregate_device_noalloc(".a", b.ptr, b.n)
        r, _ = w.aggregate_device_noalloc("count", b.ptr, b.n)
        assert r.agg.count == 1
    finally:
        b.free()
