"""Host side of the display path: the gamma map of post-process-shader.frag
:62-68 as the library tabulates it, and the PNG/PPM image writer."""
import struct
import zlib

import numpy as np
import pytest


def test_gamma_table_is_pow_1_over_2_2(rt):
    t = rt.gamma_table().astype(np.int64)
    assert t[0] == 0 and t[255] == 255
    assert (np.diff(t) >= 0).all()
    exact = 255.0 * (np.arange(256) / 255.0) ** (1.0 / 2.2)       # double precision, no float32 steps
    assert (np.abs(t - exact) <= 0.5 + 1e-3).all()
    clear = np.abs(exact - np.floor(exact) - 0.5) > 1e-3            # away from a rounding tie
    assert np.array_equal(t[clear], np.rint(exact[clear]).astype(np.int64))
    assert t[128] == 186 and t[1] == 21 and t[64] == 136               # values of 255*(c/255)^(1/2.2)


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xffffffff, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])                               # checks the Adler-32 too
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 4)


@pytest.mark.parametrize("w,h", [(1, 1), (7, 5), (300, 250)])        # 300x250 spans several stored blocks
def test_png_round_trip(rt, tmp_path, w, h):
    img = np.random.default_rng(w * h).integers(0, 256, (h, w, 4), dtype=np.uint8)
    p = str(tmp_path / "img.png")
    rt.image_write(p, img)
    assert np.array_equal(read_png(p), img)


def test_ppm_round_trip(rt, tmp_path):
    img = np.random.default_rng(3).integers(0, 256, (9, 13, 4), dtype=np.uint8)
    p = str(tmp_path / "img.ppm")
    rt.image_write(p, img)
    data = open(p, "rb").read()
    head = b"P6\n13 9\n255\n"
    assert data[:len(head)] == head
    assert np.array_equal(np.frombuffer(data[len(head):], np.uint8).reshape(9, 13, 3), img[..., :3])


def test_image_write_rejects_unknown_extension(rt, tmp_path):
    with pytest.raises(rt.RtError):
        rt.image_write(str(tmp_path / "img.bmp"), np.zeros((2, 2, 4), np.uint8))
