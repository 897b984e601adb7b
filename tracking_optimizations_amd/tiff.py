"""Minimal baseline-TIFF reader for ``batch_file`` (the reference decodes TIFF
pages with the `tiff` crate, decoderiter.rs:9-71, converting u8/u16 to f32).

Supports uncompressed (compression 1), single-sample, 8- or 16-bit unsigned
grey pages in strips, little- or big-endian, any number of pages of one
shape.  Other encodings raise NotImplementedError.
"""
from __future__ import annotations

import struct

import numpy as np

_TYPES = {1: ("B", 1), 3: ("H", 2), 4: ("I", 4), 16: ("Q", 8)}


def _read_ifd(buf: bytes, off: int, e: str):
    (n,) = struct.unpack_from(e + "H", buf, off)
    tags = {}
    for i in range(n):
        tag, typ, cnt, val = struct.unpack_from(e + "HHI4s", buf, off + 2 + 12 * i)
        if typ not in _TYPES:
            continue
        fmt, size = _TYPES[typ]
        if cnt * size <= 4:
            vals = struct.unpack_from(e + fmt * cnt, val)
        else:
            (ptr,) = struct.unpack_from(e + "I", val)
            vals = struct.unpack_from(e + fmt * cnt, buf, ptr)
        tags[tag] = vals
    (nxt,) = struct.unpack_from(e + "I", buf, off + 2 + 12 * n)
    return tags, nxt


def read_tiff(path: str) -> np.ndarray:
    """All pages as an (n_pages, height, width) uint8/uint16 array."""
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:2] == b"II":
        e = "<"
    elif buf[:2] == b"MM":
        e = ">"
    else:
        raise ValueError(f"{path}: not a TIFF file")
    magic, off = struct.unpack_from(e + "HI", buf, 2)
    if magic != 42:
        raise NotImplementedError(f"{path}: BigTIFF is not supported")
    pages = []
    while off:
        tags, off = _read_ifd(buf, off, e)
        w, h = tags[256][0], tags[257][0]
        bps = tags.get(258, (1,))[0]
        comp = tags.get(259, (1,))[0]
        spp = tags.get(277, (1,))[0]
        fmt = tags.get(339, (1,))[0]
        if comp != 1 or spp != 1 or bps not in (8, 16) or fmt != 1:
            raise NotImplementedError(
                f"{path}: only uncompressed 8/16-bit unsigned grey pages are supported")
        dt = np.dtype(np.uint8) if bps == 8 else np.dtype(e + "u2")
        offs, counts = tags[273], tags[279]
        data = b"".join(buf[o:o + c] for o, c in zip(offs, counts))
        page = np.frombuffer(data, dt, count=w * h).reshape(h, w)
        pages.append(page.astype(np.uint8 if bps == 8 else np.uint16))
    if not pages:
        raise ValueError(f"{path}: no pages")
    if len({p.shape for p in pages}) != 1:
        raise NotImplementedError(f"{path}: pages of different shapes")
    return np.ascontiguousarray(np.stack(pages))
