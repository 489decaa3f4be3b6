"""Image file writers for the texture tests, from the PNG specification (ISO 15948: scanline
filters, sample packing, Adam7 interlace; deflate by Python's zlib) and the Radiance RGBE format
(Ward, Graphics Gems II: shared exponent, adaptive run-length scanlines)."""
import struct
import zlib

import numpy as np

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def paeth(a, b, c):
    """PNG spec 9.4: the neighbour closest to a + b - c, ties to a, then b."""
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    if pb <= pc:
        return b
    return c


def _pack_row(samples, depth):
    """One scanline's samples (ints) packed big-endian at `depth` bits, rows padded to a byte."""
    if depth == 8:
        return bytes(int(s) for s in samples)
    if depth == 16:
        return b"".join(struct.pack(">H", int(s)) for s in samples)
    out, acc, n = bytearray(), 0, 0
    for s in samples:
        acc = (acc << depth) | int(s)
        n += depth
        if n == 8:
            out.append(acc)
            acc, n = 0, 0
    if n:
        out.append(acc << (8 - n))
    return bytes(out)


def _filter_row(ft, row, prev, bpp):
    out = bytearray(len(row))
    for i, x in enumerate(row):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        pred = [0, a, b, (a + b) // 2, paeth(a, b, c)][ft]
        out[i] = (x - pred) & 255
    return bytes([ft]) + bytes(out)


def _chunk(tag, data):
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def png_bytes(samples, ctype, depth, interlace=0, palette=None, filters=(0, 1, 2, 3, 4), idat_parts=1, extra=()):
    """samples: (H, W, C) ints for colour type ctype at bit depth depth."""
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    raw = bytearray()
    k = 0
    for x0, y0, sx, sy in passes:
        sub = samples[y0::sy, x0::sx]
        if sub.size == 0:
            continue
        prev = bytes(len(_pack_row(sub[0].reshape(-1), depth)))
        for r in range(sub.shape[0]):
            row = _pack_row(sub[r].reshape(-1), depth)
            raw += _filter_row(filters[k % len(filters)], row, prev, bpp)
            prev = row
            k += 1
    z = zlib.compress(bytes(raw), 9)
    cut = [len(z) * i // idat_parts for i in range(idat_parts + 1)]
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace))
    for tag, data in extra:
        out += _chunk(tag, data)
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    for i in range(idat_parts):
        out += _chunk(b"IDAT", z[cut[i]:cut[i + 1]])
    return out + _chunk(b"IEND", b"")


def rgbe_encode(rgb):
    """Float RGB (H, W, 3) to RGBE bytes (H, W, 4): shared exponent of the largest component,
    mantissas floor(c * 256 / 2^e) (Ward, Graphics Gems II)."""
    v = rgb.max(axis=2)
    m, e = np.frexp(v)
    scale = np.where(v > 1e-32, np.ldexp(1.0, 8 - e.astype(np.int64)), 0.0)
    out = np.zeros(rgb.shape[:2] + (4,), np.uint8)
    out[..., :3] = np.floor(rgb * scale[..., None]).clip(0, 255).astype(np.uint8)
    out[..., 3] = np.where(v > 1e-32, e + 128, 0).astype(np.uint8)
    return out


def rgbe_decode(rgbe):
    """The format's value: mantissa * 2^(e - 136), 0 when e = 0."""
    e = rgbe[..., 3].astype(np.int64)
    f = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return (rgbe[..., :3].astype(np.float64) * f[..., None]).astype(np.float32)


def _rle_channel(vals):
    out, i, n = bytearray(), 0, len(vals)
    while i < n:
        j = i
        while j < n and j - i < 127 and vals[j] == vals[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + j - i, vals[i]])
            i = j
            continue
        j = i
        while j < n and j - i < 128 and not (j + 2 < n and vals[j] == vals[j + 1] == vals[j + 2]):
            j += 1
        j = max(j, i + 1)
        out += bytes([j - i]) + bytes(vals[i:j])
        i = j
    return bytes(out)


def hdr_bytes(rgbe, rle=True, magic=b"#?RADIANCE", fmt=b"FORMAT=32-bit_rle_rgbe", orient=None):
    h, w, _ = rgbe.shape
    head = magic + b"\n# written by a test\n" + (fmt + b"\n" if fmt else b"") + b"\n"
    head += orient if orient is not None else b"-Y %d +X %d\n" % (h, w)
    if not rle:
        return head + rgbe.tobytes()
    body = bytearray()
    for y in range(h):
        body += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            body += _rle_channel(list(rgbe[y, :, c]))
    return head + bytes(body)
