"""Static (augmented-reality) collider set: depth images to scene points and normals, for the
engine's static contact stage (arsim_engine_set_static).

The reference reads its depth maps with OpenCV (Particles/loadRangeDataFromFile.cpp:101-141,
cv::imread(path, CV_16UC1)); `read_png16` reads the same 16-bit greyscale PNG files with the
standard library (zlib), and `points_from_depth` runs the device loader (arsim_range_image_load).
"""
import ctypes as C
import struct
import zlib

import numpy as np

from . import _capi as capi


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))


def read_png16(path):
    """A non-interlaced 16-bit greyscale PNG as a (height, width) uint16 array."""
    data = open(path, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, hdr = 8, [], None
    while pos < len(data):
        (length,) = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + length]
        pos += 12 + length
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
    if hdr is None:
        raise ValueError(f"{path}: no IHDR chunk")
    w, h, depth, color, _, _, interlace = hdr
    if depth != 16 or color != 0 or interlace != 0:
        raise ValueError(f"{path}: need 16-bit greyscale, non-interlaced (got depth {depth}, colour type {color})")
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8)
    stride, bpp = 2 * w, 2
    if raw.size != h * (stride + 1):
        raise ValueError(f"{path}: image data has the wrong size")
    rows = raw.reshape(h, stride + 1)
    out = np.zeros((h, stride), np.int32)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f, line = int(rows[y, 0]), rows[y, 1:].astype(np.int32)
        if f == 0:
            cur = line
        elif f == 2:
            cur = (line + prev) & 0xFF
        elif f in (1, 3, 4):
            cur = np.zeros(stride, np.int32)
            for x in range(stride):   # depends on the pixel to the left: sequential
                a = cur[x - bpp] if x >= bpp else 0
                c = prev[x - bpp] if x >= bpp else 0
                b = prev[x]
                if f == 1:
                    pred = a
                elif f == 3:
                    pred = (a + b) >> 1
                else:
                    pred = int(_paeth(np.int32(a), np.int32(b), np.int32(c)))
                cur[x] = (line[x] + pred) & 0xFF
        else:
            raise ValueError(f"{path}: bad filter type {f} on row {y}")
        out[y] = cur
        prev = cur
    b = out.astype(np.uint8).reshape(h, w, 2)
    return (b[..., 0].astype(np.uint16) << 8) | b[..., 1].astype(np.uint16)


def write_png16(path, img, filters=None):
    """Writes a (height, width) uint16 array as a 16-bit greyscale PNG; `filters` (per row, 0..4)
    chooses the scanline filters (0 for every row by default)."""
    img = np.ascontiguousarray(img, np.uint16)
    h, w = img.shape
    be = np.zeros((h, w, 2), np.uint8)
    be[..., 0], be[..., 1] = img >> 8, img & 0xFF
    rows = be.reshape(h, 2 * w).astype(np.int32)
    out = bytearray()
    prev = np.zeros(2 * w, np.int32)
    for y in range(h):
        f = 0 if filters is None else int(filters[y])
        cur = rows[y]
        left = np.concatenate([[0, 0], cur[:-2]])
        upleft = np.concatenate([[0, 0], prev[:-2]])
        pred = {0: 0 * cur, 1: left, 2: prev, 3: (left + prev) >> 1, 4: _paeth(left, prev, upleft)}[f]
        out.append(f)
        out += ((cur - pred) & 0xFF).astype(np.uint8).tobytes()
        prev = cur

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 16, 0, 0, 0, 0))
    png += chunk(b"IDAT", zlib.compress(bytes(out))) + chunk(b"IEND", b"")
    open(path, "wb").write(png)


def _alloc(nbytes):
    p = C.c_void_p()
    capi.check(capi.lib.arsim_malloc(C.byref(p), max(int(nbytes), 4)), "malloc")
    return p


def points_from_depth(depth, camera=None, radius=1.0 / 64.0, valid_only=True):
    """Scene points and normals (float4 each) from a uint16 depth image with the device loader.

    camera: 4x4 column-major camera transformation (identity by default).  valid_only drops the
    pixels the loader leaves without a point or a normal (the image border, the pixels past
    (W-1)(H-1), and any normal that is not finite) and the pixels without depth (placed at z = -1000
    by the loader)."""
    depth = np.ascontiguousarray(depth, np.uint16)
    h, w = depth.shape
    cam = np.eye(4, dtype=np.float32) if camera is None else np.ascontiguousarray(camera, np.float32).reshape(16)
    cam = np.ascontiguousarray(cam.reshape(16))
    n = h * w
    pos = np.full((n, 4), np.nan, np.float32)
    norm = np.full((n, 4), np.nan, np.float32)
    bufs = [_alloc(depth.nbytes), _alloc(pos.nbytes), _alloc(norm.nbytes)]
    try:
        d_img, d_pos, d_norm = bufs
        capi.check(capi.lib.arsim_memcpy_htod(d_img, capi.ptr(depth), depth.nbytes, None), "htod")
        capi.check(capi.lib.arsim_memcpy_htod(d_pos, capi.ptr(pos), pos.nbytes, None), "htod")
        capi.check(capi.lib.arsim_memcpy_htod(d_norm, capi.ptr(norm), norm.nbytes, None), "htod")
        capi.check(capi.lib.arsim_range_image_load(d_img, w, h, capi.ptr(cam), float(radius), d_pos, d_norm, None,
                                                   None, None), "range_image_load")
        capi.check(capi.lib.arsim_memcpy_dtoh(capi.ptr(pos), d_pos, pos.nbytes, None), "dtoh")
        capi.check(capi.lib.arsim_memcpy_dtoh(capi.ptr(norm), d_norm, norm.nbytes, None), "dtoh")
    finally:
        for b in bufs:
            capi.lib.arsim_free(b)
    if valid_only:
        keep = np.isfinite(pos).all(1) & np.isfinite(norm).all(1) & (np.abs(pos[:, :3]).max(1) < 500.0)
        pos, norm = pos[keep], norm[keep]
    return pos, norm
