"""The embedding writer (IO::output's lines, io/IO.cpp:207-213: "(%u, %u) "
per query vertex, then a newline) formatted on the device.  The expected
text is Python's own "%u" formatting of the same rows."""
import os
import subprocess

import numpy as np
import pytest

from gpsm_amd import _lib
from gpsm_amd import graphs as G

pytestmark = pytest.mark.gpu


def expected_text(t, keep=None):
    rows = range(len(t)) if keep is None else np.nonzero(keep)[0]
    return "".join("".join("(%u, %u) " % (j, v) for j, v in enumerate(t[i])) + "\n" for i in rows).encode()


def format_tuples(ctx, t, keep=None, cap=None):
    t = np.ascontiguousarray(t, dtype=np.uint32)
    n, q = t.shape
    k = None if keep is None else np.ascontiguousarray(keep, dtype=np.uint8)
    need = C_u64()
    if cap is None:
        _lib.lib().gpsm_format_tuples(ctx.h, _lib.ptr(t), n, q, _lib.ptr(k), None, 0, need)
        cap = need.value
    out = np.zeros(max(cap, 1), dtype=np.uint8)
    rc = _lib.lib().gpsm_format_tuples(ctx.h, _lib.ptr(t), n, q, _lib.ptr(k), _lib.ptr(out), cap, need)
    return rc, out[:min(cap, need.value)].tobytes(), need.value


def C_u64():
    import ctypes
    return ctypes.c_uint64()


def values(rng, n, q):
    """Every decimal width from 1 to 10 digits, 0 and 2^32 - 1 included."""
    t = rng.integers(0, 1 << 32, size=(n, q), dtype=np.uint64)
    width = rng.integers(1, 11, size=(n, q))
    t = (t % (10 ** width)).astype(np.uint32)
    if n:
        t.flat[0] = 0
        t.flat[-1] = 0xFFFFFFFF
    return t


@pytest.mark.parametrize("q", [1, 2, 5, 11, 17, 64])
@pytest.mark.parametrize("n", [0, 1, 255, 256, 257, 3001])
def test_format_equals_python(ctx, q, n):
    rng = np.random.default_rng(q * 7919 + n)
    t = values(rng, n, q)
    rc, text, need = format_tuples(ctx, t)
    assert rc == 0 and text == expected_text(t) and need == len(text)
    keep = rng.integers(0, 2, size=n).astype(np.uint8)
    rc, text, need = format_tuples(ctx, t, keep)
    assert rc == 0 and text == expected_text(t, keep)


def test_overflow_keeps_a_prefix(ctx):
    t = values(np.random.default_rng(5), 500, 4)
    full = expected_text(t)
    rc, text, need = format_tuples(ctx, t, cap=len(full) // 3)
    assert rc == -5 and need == len(full) and text == full[:len(full) // 3]  # GPSM_EOVERFLOW


def test_many_chunks_and_fd(ctx, tmp_path):
    """260 000 rows of 64 columns cross the writer's chunk boundary: every
    line has the length Python's formatting gives it, rows around the
    boundary and a random sample are exact, and the file-descriptor writer
    produces the same bytes."""
    rng = np.random.default_rng(11)
    n, q = 260_000, 64
    t = values(rng, n, q)
    rc, text, need = format_tuples(ctx, t)
    assert rc == 0 and need == len(text)
    digits = 1 + sum((t >= 10 ** p).astype(np.int64) for p in range(1, 10))
    fixed = sum(5 + len(str(j)) for j in range(q)) + 1
    lens = digits.sum(axis=1) + fixed
    b = np.frombuffer(text, dtype=np.uint8)
    ends = np.nonzero(b == ord("\n"))[0]
    assert len(ends) == n and ends[-1] == len(b) - 1
    assert np.array_equal(np.diff(np.concatenate([[-1], ends])), lens)
    starts = np.concatenate([[0], ends[:-1] + 1])
    R = 61440 // (fixed + 10 * q)
    chunk = ((256 << 20) // (fixed + 10 * q)) // R * R
    assert chunk < n
    for i in list(range(chunk - 3, chunk + 3)) + rng.integers(0, n, size=200).tolist():
        assert text[starts[i]:ends[i] + 1] == expected_text(t[i:i + 1]), i
    path = tmp_path / "rows.txt"
    fd = os.open(str(path), os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
    got = C_u64()
    try:
        assert _lib.lib().gpsm_write_tuples(ctx.h, fd, _lib.ptr(t), n, q, None, got) == 0
    finally:
        os.close(fd)
    assert got.value == len(text) and path.read_bytes() == text


def test_cli_writes_the_reference_format(tmp_path):
    """GpSM's answer file: per query the header lines, one line per embedding
    (device-formatted), then three newlines; the embeddings are the
    enumerator's, sorted."""
    import pyoracle as O
    g = G.small_graph(31, n=400, m=1600, labels=2, kind="er")
    qs = G.queries(g, 3, 3, 4, _lib.Q_WALK_HALF, 8, 10 ** 6)
    data, query, ans = tmp_path / "d.g", tmp_path / "q.g", tmp_path / "ans.txt"
    src, dst = g.edges()
    _lib.lib().gpsm_io_write_graph(str(data).encode(), 0, 0, g.n, _lib.ptr(g.labels), len(src), _lib.ptr(src),
                                   _lib.ptr(dst))
    _lib.lib().gpsm_io_write_end(str(data).encode())
    for i, q in enumerate(qs):
        q.struct()
        _lib.lib().gpsm_io_write_graph(str(query).encode(), i, i, q.n, _lib.ptr(q.labels), q.m, _lib.ptr(q.src),
                                       _lib.ptr(q.dst))
    _lib.lib().gpsm_io_write_end(str(query).encode())
    exe = os.path.join(os.path.dirname(_lib.LIB_PATH), "GpSM")
    r = subprocess.run([exe, str(data), str(query), str(ans)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    want = b""
    for i, q in enumerate(qs):
        ref = O.o1_enumerate(g, q, tuple_cap=1 << 20)
        assert len(ref["tuples"]) == ref["count"]
        tup = np.array(sorted(map(tuple, ref["tuples"].tolist())), dtype=np.uint32).reshape(-1, q.n)
        want += b"query graph:%d    data graph:0\n%s\n" % (i, b"=" * 60) + expected_text(tup) + b"\n\n\n"
    assert ans.read_bytes() == want
