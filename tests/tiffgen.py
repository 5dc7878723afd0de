"""Writer of small complex int16 TIFF files for the reader's tests: classic
or BigTIFF, either byte order, strips (any rows per strip, stored in order
or shuffled) or tiles, and both encodings of a CInt16 pixel.  Written from
the TIFF 6.0 / BigTIFF layout, independently of the reader."""

import struct

import numpy as np


def write_ci16_tiff(path, data, order="<", big=False, rps=1, tile=None, encoding="complex",
                    shuffle=False, compression=1, sample_format=None, seed=0):
    """data: int16 (h, w, 2).  tile: (tw, th) multiples of 16 or None."""
    data = np.asarray(data, np.int16)
    h, w, _ = data.shape
    px = data.astype(order + "i2")
    chunks = []
    if tile is None:
        for s in range(0, h, rps):
            chunks.append(px[s:s + rps].tobytes())
    else:
        tw, th = tile
        for r in range(0, h, th):
            for c in range(0, w, tw):
                t = np.zeros((th, tw, 2), order + "i2")
                blk = px[r:r + th, c:c + tw]
                t[:blk.shape[0], :blk.shape[1]] = blk
                chunks.append(t.tobytes())
    hdr = 16 if big else 8
    perm = np.arange(len(chunks))
    if shuffle:
        perm = np.random.default_rng(seed).permutation(len(chunks))
    offsets = [0] * len(chunks)
    body = b""
    pos = hdr
    for k in perm:
        offsets[k] = pos
        body += chunks[k]
        pos += len(chunks[k])
    counts = [len(c) for c in chunks]
    off_type = 16 if big else 4
    if encoding == "complex":
        spp, bits, fmt = 1, [32], [5 if sample_format is None else sample_format]
    else:
        spp, bits, fmt = 2, [16, 16], [2 if sample_format is None else sample_format] * 2
    tags = [(256, 4, [w]), (257, 4, [h]), (258, 3, bits), (259, 3, [compression]),
            (262, 3, [1]), (277, 3, [spp]), (284, 3, [1]), (339, 3, fmt)]
    if tile is None:
        tags += [(273, off_type, offsets), (278, 4, [rps]), (279, 4, counts)]
    else:
        tags += [(322, 3, [tile[0]]), (323, 3, [tile[1]]), (324, off_type, offsets),
                 (325, 4, counts)]
    tags.sort()
    size = {3: 2, 4: 4, 16: 8}
    code = {3: "H", 4: "I", 16: "Q"}
    ent = 20 if big else 12
    inline = 8 if big else 4
    ifd_off = pos
    ifd_len = (8 if big else 2) + ent * len(tags) + (8 if big else 4)
    extra = b""
    extra_pos = ifd_off + ifd_len
    ifd = struct.pack(order + ("Q" if big else "H"), len(tags))
    for tag, typ, vals in tags:
        raw = struct.pack(order + code[typ] * len(vals), *vals)
        if len(raw) <= inline:
            field = raw + b"\0" * (inline - len(raw))
        else:
            field = struct.pack(order + ("Q" if big else "I"), extra_pos + len(extra))
            extra += raw
        ifd += struct.pack(order + ("HHQ" if big else "HHI"), tag, typ, len(vals)) + field
    ifd += b"\0" * (8 if big else 4)
    bo = b"II" if order == "<" else b"MM"
    if big:
        head = bo + struct.pack(order + "HHHQ", 43, 8, 0, ifd_off)
    else:
        head = bo + struct.pack(order + "HI", 42, ifd_off)
    blob = head + body + ifd + extra
    if path is None:
        return blob
    with open(path, "wb") as f:
        f.write(blob)
    return path
