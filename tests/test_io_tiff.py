"""TIFF stacks of the eSRRF command line (nanopyx_amd.io): round trips, and both directions checked against an
independent reader/writer (Pillow) and against hand-built big-endian, multi-strip files."""
import struct

import numpy as np
import pytest

from nanopyx_amd.io import imread, imwrite

PIL = pytest.importorskip("PIL.Image")


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32, np.float64, np.int32])
@pytest.mark.parametrize("shape", [(7, 5), (1, 6, 4), (5, 9, 13)])
def test_round_trip(tmp_path, dtype, shape):
    rng = np.random.default_rng(3)
    a = (rng.random(shape) * 1000).astype(dtype)
    p = tmp_path / "a.tif"
    imwrite(p, a)
    b = imread(p)
    expect = a[0] if len(shape) == 3 and shape[0] == 1 else a
    assert b.dtype == a.dtype and b.shape == expect.shape
    assert np.array_equal(b, expect)


@pytest.mark.parametrize("dtype,mode", [(np.uint16, "I;16"), (np.float32, "F"), (np.uint8, "L")])
def test_pillow_reads_ours(tmp_path, dtype, mode):
    from PIL import ImageSequence
    a = (np.arange(3 * 6 * 5).reshape(3, 6, 5) * 7).astype(dtype)
    p = tmp_path / "ours.tif"
    imwrite(p, a)
    im = PIL.open(p)
    pages = [np.array(fr) for fr in ImageSequence.Iterator(im)]
    assert len(pages) == 3 and im.mode == mode
    for k in range(3):
        assert np.array_equal(pages[k].astype(dtype), a[k])


@pytest.mark.parametrize("dtype", [np.uint16, np.float32, np.uint8])
def test_we_read_pillow(tmp_path, dtype):
    a = (np.arange(4 * 9 * 11).reshape(4, 9, 11) * 3).astype(dtype)
    frames = [PIL.fromarray(a[k]) for k in range(4)]
    p = tmp_path / "pil.tif"
    frames[0].save(p, save_all=True, append_images=frames[1:])
    b = imread(p)
    assert b.shape == a.shape and np.array_equal(b.astype(dtype), a)


def test_big_endian_multi_strip(tmp_path):
    """A hand-built big-endian file with 2 pages of 3 strips each (uint16)."""
    a = np.arange(2 * 6 * 4, dtype=np.uint16).reshape(2, 6, 4) * 11
    out = bytearray(b"MM" + struct.pack(">HI", 42, 8))
    pos = 8
    for p in range(2):
        ntag = 9
        ifd_size = 2 + 12 * ntag + 4
        data_off = pos + ifd_size + 24  # 3 offsets + 3 counts, 4 bytes each
        strip_offs = [data_off + 16 * s for s in range(3)]
        nxt = data_off + 48 if p == 0 else 0
        tags = [(256, 3, 1, 4), (257, 3, 1, 6), (258, 3, 1, 16), (259, 3, 1, 1), (262, 3, 1, 1),
                (273, 4, 3, pos + ifd_size), (277, 3, 1, 1), (278, 3, 1, 2), (279, 4, 3, pos + ifd_size + 12)]
        ifd = bytearray(struct.pack(">H", ntag))
        for tag, typ, cnt, val in tags:
            ifd += struct.pack(">HHI", tag, typ, cnt)
            ifd += struct.pack(">H2x", val) if (typ == 3 and cnt == 1) else struct.pack(">I", val)
        ifd += struct.pack(">I", nxt)
        out += ifd + struct.pack(">3I", *strip_offs) + struct.pack(">3I", 16, 16, 16)
        out += a[p].astype(">u2").tobytes()
        pos = len(out)
    path = tmp_path / "be.tif"
    path.write_bytes(bytes(out))
    b = imread(path)
    assert b.dtype == np.uint16 and np.array_equal(b, a)


def test_rejects_non_tiff(tmp_path):
    p = tmp_path / "x.tif"
    p.write_bytes(b"not a tiff")
    with pytest.raises(ValueError):
        imread(p)
