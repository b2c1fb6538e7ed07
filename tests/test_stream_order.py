"""Stream ordering of the library's device transfers (CPU: reads the
sources).  Every handle's streams are created non-blocking, so a transfer or
fill on the null stream is not ordered after work queued on them: a plain
hipMemcpy/hipMemset could land before (or after) the kernels it was meant to
follow.  The library therefore issues every copy, fill and launch on a named
stream; this test fails if product code calls a synchronous null-stream
transfer or passes the null stream to an asynchronous one."""
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = sorted(glob.glob(os.path.join(ROOT, "vring_amd", "csrc", "*.cpp")) +
                 glob.glob(os.path.join(ROOT, "vring_amd", "csrc", "*.hip")) +
                 glob.glob(os.path.join(ROOT, "vring_amd", "csrc", "*.h")))

# synchronous transfers and fills that run on the null stream
SYNC_CALLS = re.compile(r"\bhip(Memcpy|Memset|Memcpy2D|Memcpy3D|MemcpyHtoD|MemcpyDtoH|MemcpyDtoD|"
                        r"MemsetD8|MemsetD16|MemsetD32|MemcpyToSymbol|MemcpyFromSymbol)\s*\(")
ASYNC_CALLS = re.compile(r"\bhip(MemcpyAsync|MemsetAsync|Memcpy2DAsync|MemcpyToSymbolAsync|"
                         r"MemcpyFromSymbolAsync|MemsetD32Async)\s*\(")
NULL_STREAM = re.compile(r",\s*(0|nullptr|NULL|hipStreamLegacy|hipStreamPerThread)\s*\)\s*$")


def _strip_comments(src):
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.sub(r"//[^\n]*", "", src)


def _calls(src, pattern):
    """(name, argument text) of each call matching `pattern`."""
    out = []
    for m in pattern.finditer(src):
        depth, i = 0, m.end() - 1
        while i < len(src):
            if src[i] == "(":
                depth += 1
            elif src[i] == ")":
                depth -= 1
                if depth == 0:
                    break
            i += 1
        out.append((m.group(0), src[m.end() - 1:i + 1]))
    return out


def test_sources_found():
    names = {os.path.basename(p) for p in SOURCES}
    assert {"vrh_api.cpp", "vrh_kernels.hip", "vrh_particles.hip", "vrh_learn.hip"} <= names


def test_no_synchronous_null_stream_transfers():
    bad = []
    for p in SOURCES:
        src = _strip_comments(open(p).read())
        for name, _ in _calls(src, SYNC_CALLS):
            bad.append((os.path.basename(p), name))
    assert not bad, bad


def test_async_transfers_name_a_stream():
    bad = []
    for p in SOURCES:
        src = _strip_comments(open(p).read())
        for name, args in _calls(src, ASYNC_CALLS):
            flat = " ".join(args.split())
            if NULL_STREAM.search(flat) or flat.count(",") < (2 if "Memset" in name else 3):
                bad.append((os.path.basename(p), name, flat))
    assert not bad, bad


def test_kernel_launches_name_a_stream():
    """hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...): the fifth
    argument is never the null stream."""
    bad, seen = [], 0
    for p in SOURCES:
        src = _strip_comments(open(p).read())
        for _, args in _calls(src, re.compile(r"\bhipLaunchKernelGGL\s*\(")):
            seen += 1
            depth, parts, cur = 0, [], ""
            # (template arguments with commas are parenthesised, as the
            # launch macro requires)
            for ch in args[1:-1]:
                if ch in "([{":
                    depth += 1
                elif ch in ")]}":
                    depth -= 1
                if ch == "," and depth == 0:
                    parts.append(cur.strip())
                    cur = ""
                else:
                    cur += ch
            parts.append(cur.strip())
            if len(parts) < 5 or parts[4] in ("0", "nullptr", "NULL", "hipStreamLegacy"):
                bad.append((os.path.basename(p), " ".join(args.split())[:80]))
    assert seen > 40, seen
    assert not bad, bad
