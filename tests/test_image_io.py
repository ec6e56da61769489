"""Image output: the PNG and EXR files are decoded here (zlib, struct) and
compared with the pixels the reference's save_image would write
(src/util.cpp:111-150): both axes flipped, PNG channels clamp(int(255*v)),
EXR RGBA half rounded to nearest even (numpy's float16 conversion)."""
import struct
import zlib

import numpy as np
import pytest

from globillum_amd import image_io


def _image(h=7, w=5, seed=0):
    rng = np.random.default_rng(seed)
    img = rng.uniform(-0.2, 1.3, (h, w, 3)).astype(np.float32)
    img[0, 0] = [np.nan, np.inf, -np.inf]
    img[1, 1] = [0.0, 1.0, 1.0 / 255.0]
    img[2, 2] = [7e-8, 65519.0, 65520.0]   # half subnormal, largest finite, overflow
    img[3, 3] = [2.0 ** -25, 3 * 2.0 ** -26, -1e-3]
    return img


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert zlib.crc32(typ + body) == crc
        chunks.setdefault(typ, b"")
        chunks[typ] += body
        pos += 12 + n
    w, h, depth, ctype = struct.unpack(">IIBB", chunks[b"IHDR"][:10])
    assert depth == 8 and ctype == 2
    raw = zlib.decompress(chunks[b"IDAT"])
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 3)


def test_png(tmp_path):
    img = _image()
    p = tmp_path / "a.png"
    image_io.save_png(p, img)
    got = _read_png(p)
    with np.errstate(invalid="ignore"):
        s = 255.0 * img.astype(np.float32)
        want = np.where(s > 0, np.minimum(np.trunc(np.nan_to_num(s, posinf=255.0)), 255), 0).astype(np.uint8)
    assert np.array_equal(got, want[::-1, ::-1])
    assert got[-1, -1].tolist() == [0, 255, 0]        # nan -> 0, inf -> 255, -inf -> 0
    assert got[-2, -2].tolist() == [0, 255, 1]


def _read_exr(path):
    b = open(path, "rb").read()
    assert struct.unpack("<II", b[:8]) == (20000630, 2)
    pos, attrs = 8, {}
    while b[pos] != 0:
        e = b.index(b"\0", pos)
        name = b[pos:e].decode()
        e2 = b.index(b"\0", e + 1)
        typ = b[e + 1:e2].decode()
        n = struct.unpack("<i", b[e2 + 1:e2 + 5])[0]
        attrs[name] = (typ, b[e2 + 5:e2 + 5 + n])
        pos = e2 + 5 + n
    pos += 1
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    assert attrs["compression"][1] == b"\0" and attrs["channels"][0] == "chlist"
    offs = struct.unpack(f"<{h}Q", b[pos:pos + 8 * h])
    out = np.zeros((h, w, 4), np.float16)  # A, B, G, R
    for y, o in enumerate(offs):
        yy, n = struct.unpack("<ii", b[o:o + 8])
        assert yy == y and n == 8 * w
        out[y] = np.frombuffer(b[o + 8:o + 8 + n], np.float16).reshape(4, w).T
    return out


def test_exr(tmp_path):
    img = _image()
    p = tmp_path / "a.exr"
    image_io.save_exr(p, img)
    got = _read_exr(p)
    with np.errstate(over="ignore"):
        want = img.astype(np.float16)[::-1, ::-1]
    assert (got[..., 0] == np.float16(1.0)).all()
    for c, k in ((1, 2), (2, 1), (3, 0)):  # B, G, R
        g, wv = got[..., c], want[..., k]
        assert np.array_equal(g.view(np.uint16)[~np.isnan(wv)], wv.view(np.uint16)[~np.isnan(wv)])
        assert np.isnan(g[np.isnan(wv)]).all()


def test_pfm_and_errors(tmp_path):
    img = _image()
    image_io.save(tmp_path / "a.pfm", img)
    with open(tmp_path / "a.pfm", "rb") as f:
        assert f.readline() == b"PF\n" and f.readline() == b"5 7\n"
    with pytest.raises(ValueError):
        image_io.save(tmp_path / "a.jpg", img)
    with pytest.raises(Exception):
        image_io.save_png(tmp_path / "no" / "such" / "dir.png", img)
