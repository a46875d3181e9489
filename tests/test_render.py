"""The snapshot PNG of Compute::drawSnapshot (Compute.cpp:416-446), decoded
here with zlib: signature, IHDR, chunk CRCs and the reference's colours (402
grey, 400 orange, anything else white)."""
import struct
import zlib

import numpy as np

COLOURS = {402: (160, 160, 160, 255), 400: (255, 100, 0, 255)}
WHITE = (255, 255, 255, 255)


def decode_png(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    at, chunks = 8, []
    while at < len(data):
        n = struct.unpack(">I", data[at:at + 4])[0]
        typ, body = data[at + 4:at + 8], data[at + 8:at + 8 + n]
        crc = struct.unpack(">I", data[at + 8 + n:at + 12 + n])[0]
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        chunks.append((typ, body))
        at += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)                       # filter type 0 on every row
    return rows[:, 1:].reshape(h, w, 4)


def colourise(mats):
    out = np.empty(mats.shape + (4,), np.uint8)
    out[...] = WHITE
    for m, c in COLOURS.items():
        out[mats == m] = c
    return out


def test_snapshot_png_colours_and_layout(sph, tmp_path):
    rng = np.random.default_rng(0)
    mats = rng.choice([0, 400, 402, 100, -5], size=(64, 64)).astype(np.int32)
    path = tmp_path / "s.png"
    sph.write_snapshot_png(mats, str(path))
    img = decode_png(path.read_bytes())
    assert img.shape == (64, 64, 4)
    assert np.array_equal(img, colourise(mats))
