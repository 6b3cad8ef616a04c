"""Writers of small scene folders and image files for the ingestion tests (test data made at
test time: a PNG encoder with chosen per-row filters and Adam7 interlacing, COLMAP binary
models, Nerfstudio / Polycam JSON, PLY point clouds)."""
import json
import os
import struct
import zlib

import numpy as np


# ------------------------------------------------------------------------------------ PNG
def _chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _filter_rows(rows, bpp, filters):
    """rows: list of bytes; filters: filter type per row. Returns filtered bytes."""
    out = bytearray()
    prev = bytes(len(rows[0])) if rows else b""
    for r, f in zip(rows, filters):
        out.append(f)
        for i in range(len(r)):
            a = r[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = {0: 0, 1: a, 2: b, 3: (a + b) >> 1, 4: _paeth(a, b, c)}[f]
            out.append((r[i] - pred) & 0xFF)
        prev = r
    return bytes(out)


def _pack_row(samples, depth):
    """samples: 1-D int array of one row's samples -> bytes at the bit depth."""
    if depth == 8:
        return bytes(np.asarray(samples, np.uint8))
    if depth == 16:
        return np.asarray(samples, ">u2").tobytes()
    out = bytearray()
    acc, nb = 0, 0
    for v in samples:
        acc = (acc << depth) | int(v)
        nb += depth
        if nb == 8:
            out.append(acc)
            acc, nb = 0, 0
    if nb:
        out.append(acc << (8 - nb))
    return bytes(out)


def write_png(path, samples, ctype, depth, palette=None, trns=None, interlace=False, seed=0):
    """samples: [h, w, channels] integer array (palette indices for ctype 3). Rows use
    filter types cycling through 0-4 (shifted per pass) so every filter is exercised."""
    samples = np.asarray(samples)
    h, w = samples.shape[:2]
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    samples = samples.reshape(h, w, ch)
    bpp = max(1, ch * depth // 8)
    passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)] \
        if interlace else [(0, 0, 1, 1)]
    raw = b""
    for k, (x0, y0, dx, dy) in enumerate(passes):
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        rows = [_pack_row(sub[y].reshape(-1), depth) for y in range(sub.shape[0])]
        raw += _filter_rows(rows, bpp, [(y + k + seed) % 5 for y in range(len(rows))])
    png = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 1 if interlace else 0))
    if palette is not None:
        png += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        png += _chunk(b"tRNS", bytes(trns))
    z = zlib.compress(raw, 6)
    png += _chunk(b"IDAT", z[: len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])  # split across two chunks
    png += _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


# --------------------------------------------------------------------------------- COLMAP
def write_colmap(folder, cams, images, points=None):
    """cams: {id: (model, w, h, params)}; images: list of (image_id, qvec wxyz, tvec, cam_id,
    name, n_points2d); points: (xyz [n,3] float64, rgb [n,3] uint8) or None."""
    os.makedirs(folder, exist_ok=True)
    with open(os.path.join(folder, "cameras.bin"), "wb") as f:
        f.write(struct.pack("<Q", len(cams)))
        for cid, (model, w, h, params) in cams.items():
            f.write(struct.pack("<IiQQ", cid, model, w, h))
            f.write(struct.pack("<%dd" % len(params), *params))
    with open(os.path.join(folder, "images.bin"), "wb") as f:
        f.write(struct.pack("<Q", len(images)))
        for iid, q, t, cid, name, npts in images:
            f.write(struct.pack("<I4d3dI", iid, *q, *t, cid))
            f.write(name.encode() + b"\0")
            f.write(struct.pack("<Q", npts))
            for k in range(npts):
                f.write(struct.pack("<ddq", 1.5 * k, 2.5 * k, -1))
    if points is not None:
        xyz, rgb = points
        with open(os.path.join(folder, "points3D.bin"), "wb") as f:
            f.write(struct.pack("<Q", len(xyz)))
            for i, (p, c) in enumerate(zip(xyz, rgb)):
                f.write(struct.pack("<Q3d3Bd", i + 1, *p, *[int(x) for x in c], 0.5))
                track = i % 3
                f.write(struct.pack("<Q", track))
                for k in range(track):
                    f.write(struct.pack("<II", k, k))


def write_ply(path, xyz, rgb=None, ascii=False, color_type="uchar", extra=True):
    """A vertex PLY (an extra float property before the colours checks the offsets)."""
    n = len(xyz)
    props = [("float", "x"), ("float", "y"), ("float", "z")]
    if extra:
        props.append(("float", "nx"))
    if rgb is not None:
        props += [(color_type, "red"), (color_type, "green"), (color_type, "blue")]
    head = ["ply", "format %s 1.0" % ("ascii" if ascii else "binary_little_endian"), "element vertex %d" % n]
    head += ["property %s %s" % p for p in props] + ["element face 0", "property list uchar int vertex_indices",
                                                      "end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        fmt = {"float": "f", "uchar": "B", "ushort": "H", "double": "d"}
        for i in range(n):
            vals = list(map(float, xyz[i]))
            if extra:
                vals.append(0.25)
            if rgb is not None:
                vals += [rgb[i][k] for k in range(3)]
            if ascii:
                f.write((" ".join(repr(v) if isinstance(v, float) else str(v) for v in vals) + "\n").encode())
            else:
                types = [fmt[t] for t, _ in props]
                f.write(struct.pack("<" + "".join(types), *[int(v) if t in "BH" else v for v, t in zip(vals, types)]))


def write_json(path, obj):
    with open(path, "w") as f:
        json.dump(obj, f)
