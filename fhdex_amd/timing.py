"""Per-kernel timing of the fast path, read from the library's HIP events.

`enable(ctx)` starts (and zeroes) the event timing; `read(ctx)` returns
{group: (launches, total_ms)}; `dominant(ctx)` returns (group, average launch ms, launches)
of the group with the largest total.
"""
from __future__ import annotations

import ctypes as C

from .lib import check, lib


def enable(ctx, on: bool = True) -> None:
    check(lib().fhdx_ctx_timing_enable(ctx.h, 1 if on else 0), "fhdx_ctx_timing_enable")


def read(ctx) -> dict:
    out = {}
    buf = C.create_string_buffer(64)
    for k in range(lib().fhdx_timing_kinds()):
        n = C.c_int64()
        ms = C.c_double()
        check(lib().fhdx_ctx_timing_read(ctx.h, k, buf, 64, C.byref(n), C.byref(ms)),
              "fhdx_ctx_timing_read")
        out[buf.value.decode()] = (n.value, ms.value)
    return out


def dominant(ctx):
    t = read(ctx)
    name, (n, ms) = max(t.items(), key=lambda kv: kv[1][1])
    if n == 0:
        return None
    return name, ms / n, n
