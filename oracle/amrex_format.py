"""Independent reader of the AMReX plotfile / VisMF checkpoint layout (test infrastructure).

This module is the checker of fhdx_write_plotfile / fhdx_write_checkpoint: only the tests use
it, the library never does.  It parses, with numpy, what the reference's own plotfile reader
reads (exec/hydro/PDF/PDF.cpp:65-130: whitespace-separated version, ncomp, variable names,
dimension, time, finest level, prob_lo, prob_hi; then the MultiFab at <dir>/Level_0/Cell) and
what ReadCheckPoint reads (src_compressible/checkpoint.cpp:154-182: title line, step, time,
statsCount, BoxArray).  The VisMF header and FAB layout (version, how, ncomp, ngrow, BoxArray,
"FabOnDisk: <file> <offset>", per-FAB minima/maxima; "FAB <real descriptor><box> <ncomp>" then
binary64 values) follow AMReX's published format; AMReX itself is not in the reference tree,
so no file written by the reference pins those details (parity unpinned beyond the reference's
own readers).
"""
from __future__ import annotations

import os
import re

import numpy as np

REAL_DESCRIPTOR_LE64 = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"
_BOX = re.compile(r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \((\d+),(\d+),(\d+)\)\)")


def parse_box(text: str):
    m = _BOX.search(text)
    if m is None:
        raise ValueError(f"no box in {text!r}")
    v = [int(x) for x in m.groups()]
    return tuple(v[0:3]), tuple(v[3:6]), tuple(v[6:9])


def read_plotfile_header(path: str) -> dict:
    """The Header fields in the order exec/hydro/PDF/PDF.cpp:73-110 extracts them (operator>>
    on whitespace), followed by the rest of AMReX's single-level layout."""
    tok = open(os.path.join(path, "Header")).read().split()
    it = iter(tok)
    h = {"version": next(it)}
    n = int(next(it))
    h["names"] = [next(it) for _ in range(n)]
    h["dim"] = int(next(it))
    h["time"] = float(next(it))
    h["finest_level"] = int(next(it))
    h["prob_lo"] = [float(next(it)) for _ in range(h["dim"])]
    h["prob_hi"] = [float(next(it)) for _ in range(h["dim"])]
    # no refinement ratios with finest_level 0; then the level-0 domain box
    rest = " ".join(it)
    lo, hi, typ = parse_box(rest)
    h["domain"] = (lo, hi)
    rest = rest[_BOX.search(rest).end():].split()
    it = iter(rest)
    h["level_steps"] = int(next(it))
    h["dx"] = [float(next(it)) for _ in range(h["dim"])]
    h["coord"] = int(next(it))
    h["bwidth"] = int(next(it))
    lev, nbox, t = int(next(it)), int(next(it)), float(next(it))
    h["level"] = (lev, nbox, t)
    h["level_step"] = int(next(it))
    h["box_extent"] = [(float(next(it)), float(next(it))) for _ in range(h["dim"])]
    h["mf_path"] = next(it)
    return h


def read_vismf(prefix: str):
    """VisMF::Read of a one-FAB MultiFab: returns (header dict, data (ncomp, nz, ny, nx))."""
    lines = open(prefix + "_H").read().split("\n")
    h = {"version": int(lines[0]), "how": int(lines[1]), "ncomp": int(lines[2]),
         "ngrow": int(lines[3])}
    assert lines[4].split() == ["(1", "0"], lines[4]
    h["box"] = parse_box(lines[5])[:2]
    assert lines[6] == ")"
    nfab = int(lines[7])
    assert nfab == 1
    tag, fname, off = lines[8].split()
    assert tag == "FabOnDisk:"
    h["fab_file"], h["offset"] = fname, int(off)
    assert lines[9] == ""
    mm = []
    q = 10
    for _ in range(2):
        m, n = (int(x) for x in lines[q].split(","))
        assert (m, n) == (1, h["ncomp"])
        vals = [float(x) for x in lines[q + 1].split(",") if x]
        assert len(vals) == h["ncomp"]
        mm.append(np.array(vals))
        assert lines[q + 2] == ""
        q += 3
    h["min"], h["max"] = mm
    with open(os.path.join(os.path.dirname(prefix), fname), "rb") as f:
        f.seek(h["offset"])
        head = f.readline().decode()
        assert head.startswith("FAB " + REAL_DESCRIPTOR_LE64), head
        body = head[len("FAB " + REAL_DESCRIPTOR_LE64):]
        lo, hi, typ = parse_box(body)
        nc = int(body[_BOX.search(body).end():].split()[0])
        assert nc == h["ncomp"]
        shape = (nc, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
        data = np.frombuffer(f.read(8 * int(np.prod(shape))), dtype="<f8").reshape(shape)
    h["fab_box"] = (lo, hi)
    return h, data


def read_checkpoint_header(path: str) -> dict:
    """ReadCheckPoint's header reads (checkpoint.cpp:154-182), plus the optional seed line."""
    lines = open(os.path.join(path, "Header")).read().split("\n")
    h = {"title": lines[0], "step": int(lines[1]), "time": float(lines[2]),
         "statsCount": int(lines[3])}
    assert lines[4].split() == ["(1", "0"], lines[4]
    h["box"] = parse_box(lines[5])[:2]
    assert lines[6] == ")"
    h["seed"] = None
    if len(lines) > 7 and lines[7].startswith("fhdx_philox_seed "):
        h["seed"] = int(lines[7].split()[1])
    return h
