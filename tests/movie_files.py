"""Independent TIFF and ETS writers for the movie-reader tests: TIFF 6.0
strips (PackBits, Deflate, LZW, horizontal predictor), BigTIFF, and the ETS
layout read by MinimalETSParser (gpu-tracking/src/decoderiter.rs:127-183)."""
import struct
import zlib

import numpy as np


def packbits(data: bytes) -> bytes:
    out, i, n = bytearray(), 0, len(data)
    while i < n:
        j = i
        while j + 1 < n and data[j + 1] == data[i] and j - i < 127:
            j += 1
        if j > i:
            out += struct.pack("b", -(j - i)) + data[i:i + 1]
            i = j + 1
        else:
            k = i
            while k < n and k - i < 128 and not (k + 1 < n and data[k + 1] == data[k]):
                k += 1
            k = max(k, i + 1)
            out += struct.pack("b", k - i - 1) + data[i:k]
            i = k
    return bytes(out)


def lzw(data: bytes) -> bytes:
    """TIFF LZW encoder: MSB-first codes, Clear first, EOI last, the width
    grows when the next free code reaches 2^width."""
    out = bytearray()
    state = {"bits": 0, "n": 0, "w": 9}

    def put(code):
        state["bits"] = (state["bits"] << state["w"]) | code
        state["n"] += state["w"]
        while state["n"] >= 8:
            state["n"] -= 8
            out.append((state["bits"] >> state["n"]) & 0xFF)
        state["bits"] &= (1 << state["n"]) - 1

    table = {bytes([i]): i for i in range(256)}
    nxt = 258
    put(256)
    w = b""
    for c in data:
        wc = w + bytes([c])
        if wc in table:
            w = wc
            continue
        put(table[w])
        table[wc] = nxt
        nxt += 1
        if nxt >= 4094:
            put(256)
            table = {bytes([i]): i for i in range(256)}
            nxt = 258
            state["w"] = 9
        elif nxt >= (1 << state["w"]):
            state["w"] += 1
        w = bytes([c])
    if w:
        put(table[w])
        nxt += 1
        if nxt >= (1 << state["w"]) and state["w"] < 12:
            state["w"] += 1
    put(257)
    if state["n"]:
        out.append((state["bits"] << (8 - state["n"])) & 0xFF)
    return bytes(out)


def write_tiff(path, pages, big_endian=False, bigtiff=False, rows_per_strip=None,
               compression=1, predictor=1):
    e = ">" if big_endian else "<"
    h, w = pages[0].shape
    bps = pages[0].dtype.itemsize * 8
    rps = rows_per_strip or h
    strips_per_page = []
    for p in pages:
        a = p.astype(p.dtype.newbyteorder("="))
        if predictor == 2:
            a = a.copy()
            a[:, 1:] = (a[:, 1:].astype(np.int64) - a[:, :-1].astype(np.int64)) % (1 << bps)
            a = a.astype(p.dtype)
        a = a.astype(p.dtype.newbyteorder(e))
        strips = []
        for r0 in range(0, h, rps):
            raw = a[r0:r0 + rps].tobytes()
            if compression == 32773:
                raw = packbits(raw)
            elif compression == 8:
                raw = zlib.compress(raw)
            elif compression == 5:
                raw = lzw(raw)
            strips.append(raw)
        strips_per_page.append(strips)
    out = bytearray((b"MM" if big_endian else b"II"))
    if bigtiff:
        out += struct.pack(e + "HHHQ", 43, 8, 0, 16)
        hdr, ent, cnt_f, off_f, nfmt = 16, 20, "Q", "Q", "Q"
    else:
        out += struct.pack(e + "HI", 42, 8)
        hdr, ent, cnt_f, off_f, nfmt = 8, 12, "I", "I", "H"
    nstrips = len(strips_per_page[0])
    ntags = 11
    ifd_size = struct.calcsize(e + nfmt) + ntags * ent + struct.calcsize(e + off_f)
    arrays = 2 * nstrips * 8
    pos = hdr
    layout = []
    for _ in pages:
        layout.append(pos)
        pos += ifd_size + arrays
    data_pos = []
    for strips in strips_per_page:
        ps = []
        for s in strips:
            ps.append(pos)
            pos += len(s)
        data_pos.append(ps)
    for i, strips in enumerate(strips_per_page):
        ifd = layout[i]
        arr_off = ifd + ifd_size
        nxt = layout[i + 1] if i + 1 < len(pages) else 0
        so, sc = data_pos[i], [len(s) for s in strips]
        word = 16 if bigtiff else 4   # LONG8 in BigTIFF, LONG otherwise
        wfmt = "Q" if bigtiff else "I"
        tags = [(256, 4, [w]), (257, 4, [h]), (258, 3, [bps]), (259, 3, [compression]),
                (262, 3, [1]), (273, word, so), (277, 3, [1]), (278, 4, [rps]),
                (279, word, sc), (284, 3, [1]), (317, 3, [predictor])]
        assert len(tags) == ntags
        blob = bytearray(struct.pack(e + nfmt, ntags))
        ext = bytearray()
        for tag, typ, vals in tags:
            fmt = {3: "H", 4: "I", 16: "Q"}[typ]
            payload = struct.pack(e + fmt * len(vals), *vals)
            inl = 8 if bigtiff else 4
            if len(payload) <= inl:
                v = payload.ljust(inl, b"\0")
            else:
                v = struct.pack(e + off_f, arr_off + len(ext)).ljust(inl, b"\0")
                ext += payload
            blob += struct.pack(e + "HH" + cnt_f, tag, typ, len(vals)) + v
        blob += struct.pack(e + off_f, nxt)
        assert len(blob) == ifd_size and len(ext) <= arrays
        out += blob + ext.ljust(arrays, b"\0")
    for strips in strips_per_page:
        for s in strips:
            out += s
    open(path, "wb").write(bytes(out))


def write_ets(path, frames, dims, chunks, dtype_code=4, ndim=5):
    """chunks: list of (channel, frame coordinate, frame array index)."""
    d0, d1 = dims
    hdr_off, table_off = 64, 128
    rec = 4 + 4 * ndim + 16
    data_off = table_off + rec * len(chunks)
    out = bytearray(b"SIS\0" + b"\0" * 8 + struct.pack("<IQ", ndim, hdr_off) + b"\0" * 8 +
                    struct.pack("<QI", table_off, len(chunks)))
    out = out.ljust(hdr_off, b"\0")
    out += b"ETS\0" + b"\0" * 4 + struct.pack("<I", dtype_code) + b"\0" * 16 + \
        struct.pack("<II", d0, d1)
    out = out.ljust(table_off, b"\0")
    for k, (ch, fr, _) in enumerate(chunks):
        coord = [0] * ndim
        coord[2], coord[4] = fr, ch
        out += b"\0" * 4 + struct.pack("<" + "I" * ndim, *coord) + \
            struct.pack("<Q", data_off + k * d0 * d1 * 2) + b"\0" * 8
    for (_, _, idx) in chunks:
        out += frames[idx].astype("<u2").tobytes()
    open(path, "wb").write(bytes(out))
