"""Every device operation of the library is issued on its own stream.

The library stream is non-blocking, so the legacy null stream is not
ordered against it: a plain hipMemset of a new vector can land after a
later copy into it on the library stream (the copy's data then reads as
zeros), and a plain hipMemcpy can overtake kernels still running.  The
sources must therefore use the stream-ordered helpers (h2d, d2h, dzero)
or explicit *Async calls on stream(); this test reads the sources."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nekbone_amd", "csrc")

NULL_STREAM = re.compile(r"\bhip(Memcpy|Memset|Memcpy2D|MemcpyToSymbol|MemcpyFromSymbol|"
                         r"MemsetD32|MemsetD8)\s*\(")


def _sources():
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".hip", ".cpp", ".h")):
            with open(os.path.join(CSRC, f)) as fh:
                yield f, fh.read()


def test_no_null_stream_copies_or_fills():
    bad = []
    for f, text in _sources():
        for ln, line in enumerate(text.splitlines(), 1):
            code = line.split("//", 1)[0]
            if NULL_STREAM.search(code):
                bad.append("%s:%d: %s" % (f, ln, line.strip()))
    assert not bad, "null-stream copy or fill:\n" + "\n".join(bad)


def test_async_calls_name_the_library_stream():
    """hipMemcpyAsync / hipMemsetAsync take stream() (never 0 / nullptr)."""
    bad = []
    call = re.compile(r"hip(MemcpyAsync|MemsetAsync)\s*\(")
    for f, text in _sources():
        for m in call.finditer(text):
            depth, i = 1, m.end()
            while depth and i < len(text):
                depth += {"(": 1, ")": -1}.get(text[i], 0)
                i += 1
            args = text[m.end():i - 1]
            if "stream()" not in args.rsplit(",", 1)[-1]:
                bad.append("%s: %s" % (f, " ".join(args.split())))
    assert not bad, "async copy or fill off the library stream:\n" + "\n".join(bad)
