"""Annotated binary (.abn) files and config files, as the reference writes
them (src/abn.c:186-283, src/cfg.c:297-312).  Used by the tests and bench.py
to write synthetic inputs (cellfile, pkfile, pixfile) and read program outputs.
"""
from __future__ import annotations

import re
import struct

import numpy as np

TAG = re.compile(rb"#!Begin data block: n = (\d+), size = (\d+), endian = (.), format = (\S+)")


def _cfg_lines(opts):
    # keys in strcmp order, doubles with %g (src/cfg.c:297-312, :346)
    out = []
    for k in sorted(opts, key=lambda s: s.encode()):
        v = opts[k]
        if isinstance(v, float):
            v = "%g" % v
        v = str(v)
        if v and (v[0].isspace() or v[-1].isspace()):
            out.append(f'{k} = "{v}"\n')
        else:
            out.append(f"{k} = {v}\n")
    return "".join(out)


def write_header(f, n, fmt, opts=None, comments=()):
    for c in comments:
        f.write(("# " + c + "\n").encode())
    f.write(b"\n")
    if opts:
        f.write(_cfg_lines(opts).encode())
    size = struct.calcsize("<" + fmt)
    f.write(f"#!Begin data block: n = {n}, size = {size}, endian = L, format = {fmt}\n".encode())


def write(path, array, fmt, opts=None, comments=()):
    a = np.ascontiguousarray(array)
    with open(path, "wb") as f:
        write_header(f, len(a), fmt, opts, comments)
        f.write(a.tobytes())


def read_blocks(path):
    """All data blocks of an .abn file: list of (opts, fmt, raw bytes)."""
    data = open(path, "rb").read()
    blocks = []
    pos = 0
    while pos < len(data):
        m = TAG.search(data, pos)
        if not m:
            break
        header = data[pos:m.start()].decode(errors="replace")
        opts = {}
        for line in header.splitlines():
            s = line.strip()
            if not s or s.startswith("#") or "=" not in s:
                continue
            k, v = s.split("=", 1)
            opts[k.strip()] = v.strip().strip('"')
        n, size = int(m.group(1)), int(m.group(2))
        start = data.index(b"\n", m.end()) + 1
        raw = data[start:start + n * size]
        blocks.append((opts, m.group(4).decode(), raw, n))
        pos = start + n * size
    return blocks


def read_doubles(path):
    """Concatenated double blocks + options of the first block."""
    blocks = read_blocks(path)
    return blocks[0][0], [np.frombuffer(b[2], dtype="<f8") for b in blocks]


def read_cells(path):
    """(options, cells) of a cellfile ("4i8d" records)."""
    from .lib import CELL_DTYPE
    opts, fmt, raw, n = read_blocks(path)[0]
    assert fmt == "4i8d", fmt
    return opts, np.frombuffer(raw, dtype=CELL_DTYPE).copy()


def write_cells(path, cells, survey_opts):
    opts = dict(survey_opts)
    opts["Ncells"] = len(cells)
    write(path, cells, "4i8d", opts)


def write_pkfile(path, k, P):
    np.savetxt(path, np.c_[k, P], fmt="%.17g")


def write_config(path, d):
    with open(path, "w") as f:
        for k, v in d.items():
            f.write(f"{k} = {v}\n")
