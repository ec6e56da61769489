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


# ---- reading: PNG and PPM texture images --------------------------------
def _png_bytes(raw_rows, w, h, depth, ctype, plte=None, trns=None, interlace=0):
    """Encode filtered scanlines (filter byte + row bytes each) as a PNG."""
    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body))
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace))
    if plte is not None:
        out += chunk(b"PLTE", bytes(plte))
    if trns is not None:
        out += chunk(b"tRNS", bytes(trns))
    return out + chunk(b"IDAT", zlib.compress(raw_rows)) + chunk(b"IEND", b"")


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def _filter_rows(rows, bpp):
    """Apply filter (row index % 5) to each row (bytes), as an encoder would."""
    out, prev = b"", bytes(len(rows[0]))
    for y, r in enumerate(rows):
        ft = y % 5
        f = bytearray()
        for i, x in enumerate(r):
            a = r[i - bpp] if i >= bpp else 0
            up = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, up, (a + up) // 2, _paeth(a, up, c)][ft]
            f.append((x - pred) & 255)
        out += bytes([ft]) + bytes(f)
        prev = r
    return out


@pytest.mark.parametrize("ctype,depth", [(2, 8), (6, 8), (0, 8), (4, 8), (2, 16), (6, 16), (0, 16), (0, 1),
                                         (0, 2), (0, 4), (3, 8), (3, 4), (3, 1)])
def test_png_reader(tmp_path, ctype, depth):
    rng = np.random.default_rng(ctype * 100 + depth)
    w, h = 13, 9
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    plte = trns = None
    if ctype == 3:
        n = 1 << depth
        plte = rng.integers(0, 256, 3 * n).astype(np.uint8)
        trns = rng.integers(0, 256, max(1, n // 2)).astype(np.uint8)
        idx = rng.integers(0, n, (h, w))
        want = np.concatenate([plte.reshape(-1, 3)[idx], np.full((h, w, 1), 255, np.uint8)], -1)
        want[..., 3] = np.where(idx < len(trns), trns[np.minimum(idx, len(trns) - 1)], 255)
        samples = idx
    else:
        maxv = (1 << depth) - 1
        samples = rng.integers(0, maxv + 1, (h, w, ch))
        v8 = (samples >> 8) if depth == 16 else (samples * 255 // maxv)
        rgb = np.repeat(v8[..., :1], 3, -1) if ch <= 2 else v8[..., :3]
        alpha = v8[..., -1:] if ch in (2, 4) else np.full((h, w, 1), 255)
        want = np.concatenate([rgb, alpha], -1).astype(np.uint8)
    rows = []
    for y in range(h):
        if depth == 16:
            rows.append(b"".join(int(s).to_bytes(2, "big") for s in samples[y].reshape(-1)))
        elif depth == 8:
            rows.append(bytes(np.asarray(samples[y], np.uint8).reshape(-1)))
        else:  # sub-byte, one channel, packed from the high bits
            bits = "".join(format(int(s), f"0{depth}b") for s in samples[y].reshape(-1))
            bits += "0" * (-len(bits) % 8)
            rows.append(bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8)))
    bpp = max(1, ch * depth // 8)
    p = tmp_path / "t.png"
    p.write_bytes(_png_bytes(_filter_rows(rows, bpp), w, h, depth, ctype, plte, trns))
    got = image_io.read_image(p)
    assert got.shape == (h, w, 4) and np.array_equal(got, want)


def test_ppm_reader_and_png_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)
    p = tmp_path / "a.ppm"
    p.write_bytes(b"P6\n# comment\n7 5\n255\n" + img.tobytes())
    got = image_io.read_image(p)
    assert np.array_equal(got[..., :3], img) and (got[..., 3] == 255).all()
    wide = rng.integers(0, 1001, (5, 7, 3))
    p.write_bytes(b"P6 7 5 1000\n" + b"".join(int(x).to_bytes(2, "big") for x in wide.reshape(-1)))
    got = image_io.read_image(p)
    assert np.array_equal(got[..., :3], ((wide * 255 + 500) // 1000).astype(np.uint8))
    # what save_png writes (both axes flipped, clamp(int(255 v))) reads back
    f = rng.uniform(0, 1, (6, 4, 3)).astype(np.float32)
    image_io.save_png(tmp_path / "r.png", f)
    back = image_io.read_image(tmp_path / "r.png")
    assert np.array_equal(back[..., :3], np.minimum((255 * f).astype(np.int32), 255).astype(np.uint8)[::-1, ::-1])


def test_reader_errors(tmp_path):
    from globillum_amd._lib import GiError
    rows = _filter_rows([bytes(12)] * 2, 3)
    p = tmp_path / "i.png"
    p.write_bytes(_png_bytes(rows, 4, 2, 8, 2, interlace=1))
    with pytest.raises(GiError):
        image_io.read_image(p)
    p.write_bytes(_png_bytes(rows, 4, 2, 8, 2)[:40])
    with pytest.raises(GiError):
        image_io.read_image(p)
    p.write_bytes(b"GIF89a....")
    with pytest.raises(GiError):
        image_io.read_image(p)
    with pytest.raises(GiError):
        image_io.read_image(tmp_path / "missing.png")
