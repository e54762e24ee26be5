"""Host-side tests of the frame writers (blz_write_image): each file is
decoded here from its format specification and compared with the image
(PNG chunks and CRCs, zlib stream, 8-bit quantisation; PPM text; EXR
header, offset table and FLOAT planes bit for bit)."""
import struct
import zlib

import numpy as np
import pytest

B = pytest.importorskip("blaze_amd")


def image(h=37, w=70, seed=0):
    rng = np.random.default_rng(seed)
    img = rng.uniform(-0.2, 1.3, (h, w, 3)).astype(np.float32)
    img[0, 0] = [0.0, 1.0, 0.5]
    return img


def test_png(tmp_path):
    img = image()
    p = str(tmp_path / "f_0001.png")
    B.write_image(p, img)
    data = open(p, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF
        chunks.setdefault(typ, b"")
        chunks[typ] += body
        pos += 12 + n
    w, h, depth, ctype = struct.unpack(">IIBB", chunks[b"IHDR"][:10])
    assert (w, h, depth, ctype) == (70, 37, 8, 6) and b"IEND" in chunks
    raw = zlib.decompress(chunks[b"IDAT"])
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert not rows[:, 0].any()
    px = rows[:, 1:].reshape(h, w, 4)
    want = np.clip((img.astype(np.float64) * 255.0).astype(np.int64), 0, 255)
    assert np.array_equal(px[..., :3], want) and (px[..., 3] == 255).all()


def test_ppm(tmp_path):
    img = image(5, 8)
    p = str(tmp_path / "f_0001.ppm")
    B.write_image(p, img)
    tok = open(p).read().split()
    assert tok[:4] == ["P3", "8", "5", "255"]
    vals = np.array(tok[4:], int).reshape(5, 8, 3)
    assert np.array_equal(vals, np.trunc(255.99 * img.astype(np.float64)).astype(int))


def test_exr(tmp_path):
    img = image(9, 13)
    p = str(tmp_path / "f_0001.exr")
    B.write_image(p, img)
    d = open(p, "rb").read()
    magic, version = struct.unpack("<ii", d[:8])
    assert magic == 20000630 and version == 2
    pos, attrs = 8, {}
    while d[pos] != 0:
        e = d.index(b"\0", pos)
        name = d[pos:e].decode()
        e2 = d.index(b"\0", e + 1)
        typ = d[e + 1:e2].decode()
        n, = struct.unpack("<i", d[e2 + 1:e2 + 5])
        attrs[name] = (typ, d[e2 + 5:e2 + 5 + n])
        pos = e2 + 5 + n
    pos += 1
    assert attrs["compression"] == ("compression", b"\0")
    assert struct.unpack("<4i", attrs["dataWindow"][1]) == (0, 0, 12, 8)
    assert attrs["channels"][1].startswith(b"B\0")
    offs = struct.unpack(f"<{9}Q", d[pos:pos + 8 * 9])
    for j, o in enumerate(offs):
        y, n = struct.unpack("<ii", d[o:o + 8])
        assert y == j and n == 13 * 3 * 4
        planes = np.frombuffer(d[o + 8:o + 8 + n], np.float32).reshape(3, 13)
        for c, k in enumerate((2, 1, 0)):   # B, G, R
            assert np.array_equal(planes[c].view(np.uint32), img[j, :, k].view(np.uint32))


def test_bad_extension(tmp_path):
    with pytest.raises(B.BlazeError, match="extension"):
        B.write_image(str(tmp_path / "f.bmp"), image(2, 2))
