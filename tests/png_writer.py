"""A small PNG encoder for tests, written from the PNG specification (W3C PNG 2nd edition): chunks
with CRC-32, zlib-compressed scanlines with a chosen filter per row (None, Sub, Up, Average, Paeth),
optional Adam7 interlacing, bit depth 8 (and 16 for error cases), every colour type. The decoded
RGB8 planes a conforming decoder must return are computed from the same pixels (`expected_rgb8`)."""
import struct
import zlib

import numpy as np

CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def chunk(kind: bytes, data: bytes, bad_crc: bool = False) -> bytes:
    crc = zlib.crc32(kind + data) & 0xFFFFFFFF
    if bad_crc:
        crc ^= 1
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", crc)


def _predict(kind, a, b, c):
    if kind == 0:
        return 0
    if kind == 1:
        return a
    if kind == 2:
        return b
    if kind == 3:
        return (a + b) // 2
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else b if pb <= pc else c


def filter_rows(img: np.ndarray, bpp: int, filters) -> bytes:
    """img: (h, row_bytes) uint8 raw scanlines of one (sub-)image."""
    out = bytearray()
    prev = np.zeros(img.shape[1], np.int64)
    for y in range(img.shape[0]):
        kind = filters if isinstance(filters, int) else (y * 7 + 3) % 5
        row = img[y].astype(np.int64)
        enc = np.zeros_like(row)
        for i in range(row.size):
            a = row[i - bpp] if i >= bpp else 0
            c = prev[i - bpp] if i >= bpp else 0
            enc[i] = (row[i] - _predict(kind, a, prev[i], c)) & 0xFF
        out.append(kind)
        out += enc.astype(np.uint8).tobytes()
        prev = row
    return bytes(out)


def png_bytes(pix: np.ndarray, colour: int, *, filters="mixed", interlace=False, palette=None, bit_depth=8,
              idat_parts=1, ancillary_bad_crc=False, idat_bad_crc=False, truncate=0) -> bytes:
    """pix: (h, w, channels) uint8 samples (uint16 for bit depth 16), indices (h, w, 1) for palette."""
    h, w, ch = pix.shape
    assert ch == CHANNELS[colour]
    f = -1 if filters == "mixed" else filters
    flt = f if f >= 0 else "mixed"
    sample_bytes = 2 if bit_depth == 16 else 1
    bpp = ch * sample_bytes

    def raw(img):
        if bit_depth == 16:
            return img.astype(">u2").view(np.uint8).reshape(img.shape[0], -1)
        return img.astype(np.uint8).reshape(img.shape[0], -1)

    if not interlace:
        data = filter_rows(raw(pix), bpp, flt if flt != "mixed" else "mixed")
    else:
        data = b""
        for x0, y0, dx, dy in ADAM7:
            sub = pix[y0::dy, x0::dx]
            if sub.shape[0] and sub.shape[1]:
                data += filter_rows(raw(sub), bpp, flt if flt != "mixed" else "mixed")
    z = zlib.compress(data, 6)
    parts = [z[i * len(z) // idat_parts:(i + 1) * len(z) // idat_parts] for i in range(idat_parts)]
    out = b"\x89PNG\r\n\x1a\n"
    out += chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, bit_depth, colour, 0, 0, 1 if interlace else 0))
    out += chunk(b"tEXt", b"Comment\x00test", bad_crc=ancillary_bad_crc)
    if palette is not None:
        out += chunk(b"PLTE", np.asarray(palette, np.uint8).tobytes())
    for i, p in enumerate(parts):
        out += chunk(b"IDAT", p, bad_crc=idat_bad_crc and i == 0)
    out += chunk(b"IEND", b"")
    return out[:len(out) - truncate] if truncate else out


def split_chunks(data: bytes):
    """[[type, body, crc], ...] of a PNG file."""
    pos, out = 8, []
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        out.append([data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n], struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]])
        pos += 12 + n
    return out


def join_chunks(chunks) -> bytes:
    return b"\x89PNG\r\n\x1a\n" + b"".join(struct.pack(">I", len(b)) + k + b + struct.pack(">I", c) for k, b, c in chunks)


def expected_rgb8(pix: np.ndarray, colour: int, palette=None) -> np.ndarray:
    """Planes (3, h, w): grey replicated, alpha dropped, palette looked up (entries past its end 0)."""
    if colour in (0, 4):
        g = pix[..., 0]
        return np.stack([g, g, g]).astype(np.uint8)
    if colour == 3:
        lut = np.zeros((256, 3), np.uint8)
        pal = np.asarray(palette, np.uint8).reshape(-1, 3)
        lut[:len(pal)] = pal
        return np.moveaxis(lut[pix[..., 0]], -1, 0).copy()
    return np.moveaxis(pix[..., :3], -1, 0).astype(np.uint8).copy()


def random_image(w, h, colour, seed, palette_size=200):
    rng = np.random.default_rng(seed)
    if colour == 3:
        pal = rng.integers(0, 256, size=(palette_size, 3), dtype=np.uint8)
        idx = rng.integers(0, palette_size, size=(h, w, 1), dtype=np.uint8)
        return idx, pal
    # smooth field plus noise so that every filter has non-trivial predictions
    yy, xx = np.mgrid[0:h, 0:w]
    base = (xx * 3 + yy * 5)[..., None] + rng.integers(0, 40, size=(h, w, CHANNELS[colour]))
    return (base % 256).astype(np.uint8), None


def cases():
    """The PNG decoding cases shared by the fixture generator and the tests."""
    out = []
    k = 0
    for (w, h) in [(1, 1), (7, 5), (33, 17), (64, 48)]:
        for colour in (0, 2, 3, 4, 6):
            for interlace in (False, True):
                for filters in ("mixed", 0, 1, 2, 3, 4):
                    if (w, h) == (64, 48) and filters != "mixed":
                        continue
                    out.append({"name": f"c{k}", "w": w, "h": h, "colour": colour, "interlace": interlace,
                                "filters": filters, "seed": 1000 + k, "variant": "plain"})
                    k += 1
    for variant in ("idat_split", "ancillary_bad_crc", "idat_bad_crc", "truncated", "wrong_width",
                    "depth16", "palette_short", "ihdr_bad_crc", "first_idat_bad_crc", "last_idat_bad_crc",
                    "iend_bad_crc", "no_iend", "unknown_critical_before_idat", "unknown_critical_after_idat",
                    "trailing_bytes"):
        colour = 3 if variant == "palette_short" else 2
        out.append({"name": f"e_{variant}", "w": 20, "h": 12, "colour": colour, "interlace": False,
                    "filters": "mixed", "seed": 77, "variant": variant})
    return out


def build_case(case):
    """(png bytes, width and height to decode at, expected planes or None when decoding must fail)."""
    w, h, colour, v = case["w"], case["h"], case["colour"], case["variant"]
    pix, pal = random_image(w, h, colour, case["seed"], palette_size=40 if v == "palette_short" else 200)
    if v == "palette_short":  # indices past the palette's end
        pix = (pix.astype(np.int64) * 6 % 256).astype(np.uint8)
    kw = dict(filters=case["filters"], interlace=case["interlace"], palette=pal)
    if v == "depth16":
        data = png_bytes(pix.astype(np.uint16) * 257, colour, bit_depth=16, **kw)
        return data, (w, h), None
    data = png_bytes(pix, colour, idat_parts=3 if v == "idat_split" else 1,
                     ancillary_bad_crc=v == "ancillary_bad_crc", idat_bad_crc=v == "idat_bad_crc",
                     truncate=30 if v == "truncated" else 0, **kw)
    if v in ("ihdr_bad_crc", "first_idat_bad_crc", "last_idat_bad_crc", "iend_bad_crc", "no_iend",
             "unknown_critical_before_idat", "unknown_critical_after_idat", "trailing_bytes"):
        chunks = split_chunks(png_bytes(pix, colour, idat_parts=3, **kw))
        kinds = [c[0] for c in chunks]
        idats = [i for i, k in enumerate(kinds) if k == b"IDAT"]
        flip = {"ihdr_bad_crc": 0, "first_idat_bad_crc": idats[0], "last_idat_bad_crc": idats[-1],
                "iend_bad_crc": len(chunks) - 1}.get(v)
        if flip is not None:
            chunks[flip][2] ^= 1
        if v == "no_iend":
            chunks = chunks[:-1]
        if v.startswith("unknown_critical"):
            chunks.insert(1 if v.endswith("before_idat") else idats[-1] + 1,
                          [b"ABCD", b"xyz", zlib.crc32(b"ABCDxyz") & 0xFFFFFFFF])
        data = join_chunks(chunks) + (b"trailing" if v == "trailing_bytes" else b"")
        fails = v in ("ihdr_bad_crc", "first_idat_bad_crc", "unknown_critical_before_idat")
        return data, (w, h), None if fails else expected_rgb8(pix, colour, pal)
    if v == "truncated":
        return data, (w, h), None
    if v == "wrong_width":
        return data, (w + 1, h), None
    return data, (w, h), expected_rgb8(pix, colour, pal)
