"""Image files written byte by byte from their format specifications, for
the reader's tests: PNG (ISO/IEC 15948: any colour type and depth, any
scanline filter, optional palette / tRNS / Adam7 flag), netpbm (P2, P3,
P5, P6) and PFM ("PF"/"Pf", either byte order, bottom row first).

Each writer takes integer samples (PNG, netpbm) or float32 values (PFM)
and returns the RGBA float32 array a reader must produce, computed from
the same samples: value / maxval in float32, alpha 1 where the file has
none.
"""
import struct
import zlib

import numpy as np


def _chunk(kind, data):
    body = kind + data
    return struct.pack(">I", len(data)) + body + struct.pack(">I", zlib.crc32(body) & 0xFFFFFFFF)


def _filter_row(f, row, prev, bpp):
    """PNG filter f applied to one scanline of bytes (section 9)"""
    out = bytearray(len(row))
    for i, x in enumerate(row):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        if f == 0:
            p = 0
        elif f == 1:
            p = a
        elif f == 2:
            p = b
        elif f == 3:
            p = (a + b) >> 1
        else:
            pp = a + b - c
            pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - c)
            p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
        out[i] = (x - p) & 0xFF
    return bytes(out)


def _pack_row(samples, depth):
    """one row of integer samples as bytes, most significant bits first"""
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in samples)
    if depth == 8:
        return bytes(int(v) for v in samples)
    per = 8 // depth
    out = bytearray()
    for i in range(0, len(samples), per):
        byte = 0
        for k in range(per):
            v = int(samples[i + k]) if i + k < len(samples) else 0
            byte |= v << (8 - depth * (k + 1))
        out.append(byte)
    return bytes(out)


def png_bytes(samples, ctype, depth, filters=(0,), palette=None, trns=None, interlace=0, crc_ok=True):
    """samples: (h, w, channels) integers (palette indices for ctype 3)"""
    s = np.asarray(samples)
    if s.ndim == 2:
        s = s[:, :, None]
    h, w, nch = s.shape
    bpp = max(1, (nch * depth + 7) // 8)
    raw = bytearray()
    prev = bytes(((w * nch * depth + 7) // 8))
    for y in range(h):
        row = _pack_row(s[y].reshape(-1), depth)
        f = filters[y % len(filters)]
        raw.append(f)
        raw += _filter_row(f, row, prev, bpp)
        prev = row
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace))
    if palette is not None:
        out += _chunk(b"PLTE", bytes(int(v) for v in np.asarray(palette).reshape(-1)))
    if trns is not None:
        out += _chunk(b"tRNS", bytes(trns))
    idat = _chunk(b"IDAT", zlib.compress(bytes(raw), 9))
    if not crc_ok:
        idat = idat[:-1] + bytes([idat[-1] ^ 0xFF])
    return out + idat + _chunk(b"IEND", b"")


def png_expected(samples, ctype, depth, palette=None, trns=None):
    s = np.asarray(samples)
    if s.ndim == 2:
        s = s[:, :, None]
    h, w, _ = s.shape
    out = np.ones((h, w, 4), np.float32)
    mx = np.float32((1 << depth) - 1)
    if ctype == 3:
        pal = np.asarray(palette).reshape(-1, 3)
        idx = s[:, :, 0]
        out[:, :, :3] = pal[idx].astype(np.float32) / np.float32(255)
        if trns is not None:
            t = np.ones(256, np.float32)
            t[:len(trns)] = np.asarray(list(trns), np.float32) / np.float32(255)
            out[:, :, 3] = t[idx]
        return out
    f = s.astype(np.float32) / mx
    if ctype in (0, 4):
        out[:, :, 0] = out[:, :, 1] = out[:, :, 2] = f[:, :, 0]
        if ctype == 4:
            out[:, :, 3] = f[:, :, 1]
        elif trns is not None:
            key = (trns[0] << 8) | trns[1]
            out[:, :, 3][s[:, :, 0] == key] = 0.0
    else:
        out[:, :, :3] = f[:, :, :3]
        if ctype == 6:
            out[:, :, 3] = f[:, :, 3]
        elif trns is not None:
            key = [(trns[2 * k] << 8) | trns[2 * k + 1] for k in range(3)]
            m = (s[:, :, 0] == key[0]) & (s[:, :, 1] == key[1]) & (s[:, :, 2] == key[2])
            out[:, :, 3][m] = 0.0
    return out


def pnm_bytes(samples, maxval, kind):
    """kind '2'/'5' grey, '3'/'6' RGB; samples (h, w) or (h, w, 3)"""
    s = np.asarray(samples)
    h, w = s.shape[:2]
    head = f"P{kind}\n# written for a test\n{w} {h}\n{maxval}\n".encode()
    flat = s.reshape(-1)
    if kind in "23":
        return head + " ".join(str(int(v)) for v in flat).encode() + b"\n"
    if maxval > 255:
        return head + b"".join(struct.pack(">H", int(v)) for v in flat)
    return head + bytes(int(v) for v in flat)


def pnm_expected(samples, maxval):
    s = np.asarray(samples)
    h, w = s.shape[:2]
    out = np.ones((h, w, 4), np.float32)
    f = s.astype(np.float32) / np.float32(maxval)
    if s.ndim == 2:
        out[:, :, 0] = out[:, :, 1] = out[:, :, 2] = f
    else:
        out[:, :, :3] = f
    return out


def pfm_bytes(values, little=True):
    """values: (h, w, 3) or (h, w) float32, top row first"""
    v = np.asarray(values, np.float32)
    grey = v.ndim == 2
    h, w = v.shape[:2]
    head = f"{'Pf' if grey else 'PF'}\n{w} {h}\n{-1.0 if little else 1.0}\n".encode()
    rows = v[::-1].astype("<f4" if little else ">f4")
    return head + rows.tobytes()


def pfm_expected(values):
    v = np.asarray(values, np.float32)
    h, w = v.shape[:2]
    out = np.ones((h, w, 4), np.float32)
    if v.ndim == 2:
        out[:, :, 0] = out[:, :, 1] = out[:, :, 2] = v
    else:
        out[:, :, :3] = v
    return out


def write(path, data):
    with open(path, "wb") as f:
        f.write(data)
    return str(path)
