"""Baseline TIFF stacks: what the eSRRF command line reads and writes.

The reference's command line (src/nanopyx/cli/superresolution.py:1-3, :90-104) uses ``tifffile.imread`` /
``tifffile.imwrite``. That package is not a dependency of this package, so this module implements the subset
of TIFF 6.0 those calls exchange for microscopy stacks: one image per page (IFD), uncompressed, one or several
strips, grayscale, 8/16/32-bit unsigned integers or 32/64-bit IEEE floats, either byte order. ``imread`` returns
(pages, H, W) for multi-page files and (H, W) for single pages, as tifffile does for plain multi-page files;
``imwrite`` writes little-endian, one strip per page. Compressed or tiled files raise ValueError; when tifffile is
importable it reads those instead.
"""
from __future__ import annotations

import struct

import numpy as np

_TYPES = {1: "B", 2: "s", 3: "H", 4: "I", 5: "II", 16: "Q"}  # BYTE, ASCII, SHORT, LONG, RATIONAL, LONG8
_SIZES = {1: 1, 2: 1, 3: 2, 4: 4, 5: 8, 16: 8}

WIDTH, LENGTH, BITS, COMPRESSION, PHOTOMETRIC = 256, 257, 258, 259, 262
STRIP_OFFSETS, SAMPLES, ROWS_PER_STRIP, STRIP_BYTES, PLANAR, SAMPLE_FORMAT = 273, 277, 278, 279, 284, 339
TILE_WIDTH = 322


def _dtype(bits, fmt, order):
    kind = {1: "u", 2: "i", 3: "f"}.get(fmt)
    if kind is None or bits not in (8, 16, 32, 64) or (kind == "f" and bits < 32):
        raise ValueError(f"unsupported TIFF sample format {fmt} with {bits} bits")
    return np.dtype(f"{order}{kind}{bits // 8}")


def _read_ifd(buf, off, order):
    (n,) = struct.unpack_from(order + "H", buf, off)
    tags = {}
    for k in range(n):
        tag, typ, count = struct.unpack_from(order + "HHI", buf, off + 2 + 12 * k)
        size = _SIZES.get(typ, 1) * count
        voff = off + 2 + 12 * k + 8
        if size > 4:
            (voff,) = struct.unpack_from(order + "I", buf, voff)
        if typ in (3, 4, 16):
            vals = struct.unpack_from(order + _TYPES[typ] * count, buf, voff)
            tags[tag] = list(vals)
    (nxt,) = struct.unpack_from(order + "I", buf, off + 2 + 12 * n)
    return tags, nxt


def imread(path):
    """Reads every page of an uncompressed grayscale TIFF into one array."""
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:2] == b"II":
        order = "<"
    elif buf[:2] == b"MM":
        order = ">"
    else:
        raise ValueError(f"{path}: not a TIFF file")
    (magic, off) = struct.unpack_from(order + "HI", buf, 2)
    if magic != 42:
        raise ValueError(f"{path}: BigTIFF and other variants are not supported")
    pages = []
    while off:
        tags, off = _read_ifd(buf, off, order)
        if tags.get(COMPRESSION, [1])[0] != 1 or TILE_WIDTH in tags or tags.get(SAMPLES, [1])[0] != 1:
            try:
                import tifffile  # compressed, tiled or multi-sample files
                return tifffile.imread(path)
            except ImportError:
                raise ValueError(f"{path}: only uncompressed, stripped, single-sample TIFF pages are supported")
        W, H = tags[WIDTH][0], tags[LENGTH][0]
        dt = _dtype(tags.get(BITS, [1])[0], tags.get(SAMPLE_FORMAT, [1])[0], order)
        offs, counts = tags[STRIP_OFFSETS], tags[STRIP_BYTES]
        data = b"".join(buf[o:o + c] for o, c in zip(offs, counts))
        if len(data) < H * W * dt.itemsize:
            raise ValueError(f"{path}: page data shorter than {H}x{W} samples")
        pages.append(np.frombuffer(data[:H * W * dt.itemsize], dtype=dt).reshape(H, W))
    if not pages:
        raise ValueError(f"{path}: no image pages")
    out = np.stack(pages) if len(pages) > 1 else pages[0]
    return out.astype(out.dtype.newbyteorder("="), copy=False)


def imwrite(path, data):
    """Writes a (H, W) or (pages, H, W) array (uint8/16/32, int8/16/32, float32/64) as an uncompressed TIFF."""
    a = np.asarray(data)
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3:
        raise ValueError("imwrite expects a 2D image or a 3D stack of 2D pages")
    kind = {"u": 1, "i": 2, "f": 3}.get(a.dtype.kind)
    if kind is None or a.dtype.itemsize not in (1, 2, 4, 8):
        raise ValueError(f"unsupported dtype {a.dtype}")
    a = np.ascontiguousarray(a, dtype=a.dtype.newbyteorder("<"))
    P, H, W = a.shape
    bits = a.dtype.itemsize * 8
    page_bytes = H * W * a.dtype.itemsize
    ntags = 10
    ifd_size = 2 + 12 * ntags + 4
    out = bytearray(b"II" + struct.pack("<HI", 42, 8))
    pos = 8
    for p in range(P):
        data_off = pos + ifd_size
        nxt = data_off + page_bytes if p + 1 < P else 0
        if nxt % 2:
            nxt += 1
        entries = [(WIDTH, 4, W), (LENGTH, 4, H), (BITS, 3, bits), (COMPRESSION, 3, 1), (PHOTOMETRIC, 3, 1),
                   (STRIP_OFFSETS, 4, data_off), (SAMPLES, 3, 1), (ROWS_PER_STRIP, 4, H),
                   (STRIP_BYTES, 4, page_bytes), (SAMPLE_FORMAT, 3, kind)]
        ifd = bytearray(struct.pack("<H", ntags))
        for tag, typ, val in entries:
            ifd += struct.pack("<HHI", tag, typ, 1)
            ifd += struct.pack("<H2x", val) if typ == 3 else struct.pack("<I", val)
        ifd += struct.pack("<I", nxt)
        out += ifd + a[p].tobytes()
        if len(out) % 2:
            out += b"\0"
        pos = len(out)
    with open(path, "wb") as f:
        f.write(out)
