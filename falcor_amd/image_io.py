"""Image files for rendered frames: OpenEXR (scanline, single part), PFM and 8-bit PNG.

Plays the role of Bitmap::saveImage / Bitmap::createFromFile (Source/Falcor/Utils/Image/Bitmap.cpp:
596-720, 300-360) and of ImageCompare's Image::loadFromFile / saveToFile
(Source/Tools/ImageCompare/ImageCompare.cpp:68-185) for the formats a path tracer's output needs.
Images are numpy arrays [H, W, C] (C = 3 or 4), top row first, as the reference keeps them in
memory (it flips rows for FreeImage's bottom-up scanlines).

OpenEXR: written with FLOAT or HALF channels, without compression (the reference's
ExportFlags::Uncompressed, with or without ExrFloat16) or with ZIP compression (16 scanlines per
chunk, zlib). Read: NO_COMPRESSION, RLE, ZIPS and ZIP scanline files with HALF, FLOAT or UINT
channels. PIZ (FreeImage's default for compressed EXR) is not implemented: reading a PIZ file
raises ValueError, and the compressed writer uses ZIP instead.
"""
from __future__ import annotations

import struct
import zlib
from pathlib import Path

import numpy as np

EXR_MAGIC = 20000630
_PIXEL_TYPES = {0: np.dtype("<u4"), 1: np.dtype("<f2"), 2: np.dtype("<f4")}
NO_COMPRESSION, RLE_COMPRESSION, ZIPS_COMPRESSION, ZIP_COMPRESSION, PIZ_COMPRESSION = 0, 1, 2, 3, 4
_LINES_PER_CHUNK = {NO_COMPRESSION: 1, RLE_COMPRESSION: 1, ZIPS_COMPRESSION: 1, ZIP_COMPRESSION: 16}


# --------------------------------------------------------------------------- EXR helpers
def _zip_predict(raw: bytes) -> bytes:
    """Byte reorder (even bytes, then odd bytes) followed by the delta predictor of the EXR ZIP codec."""
    a = np.frombuffer(raw, np.uint8)
    t = np.concatenate([a[0::2], a[1::2]])
    d = t.astype(np.int16)
    d[1:] = (t[1:].astype(np.int16) - t[:-1].astype(np.int16) + 128) & 0xFF
    return d.astype(np.uint8).tobytes()


def _zip_unpredict(data: bytes) -> bytes:
    d = np.frombuffer(data, np.uint8).astype(np.int64)
    if d.size == 0:
        return b""
    # t[i] = t[i-1] + d[i] - 128 (mod 256), t[0] = d[0]
    inc = d.copy()
    inc[1:] -= 128
    t = (np.cumsum(inc) & 0xFF).astype(np.uint8)
    half = (t.size + 1) // 2
    out = np.empty_like(t)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def _rle_decode(data: bytes, size: int) -> bytes:
    out = bytearray()
    i = 0
    while i < len(data) and len(out) < size:
        n = struct.unpack("b", data[i:i + 1])[0]
        i += 1
        if n < 0:
            out += data[i:i - n]
            i += -n
        else:
            out += data[i:i + 1] * (n + 1)
            i += 1
    return _zip_unpredict(bytes(out))


def _attr(name: str, typ: str, value: bytes) -> bytes:
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(value)) + value


def write_exr(path, img, half: bool = False, compression: str = "none"):
    """Writes [H, W, 3|4] (or [H, W]) as a scanline OpenEXR file with channels R, G, B (, A).

    compression: "none" (NO_COMPRESSION) or "zip" (ZIP, 16 lines per chunk)."""
    a = np.asarray(img)
    if a.ndim == 2:
        a = a[..., None]
    h, w, c = a.shape
    names = {1: ["Y"], 3: ["R", "G", "B"], 4: ["R", "G", "B", "A"]}.get(c)
    if names is None:
        raise ValueError("EXR images have 1, 3 or 4 channels")
    comp = {"none": NO_COMPRESSION, "zip": ZIP_COMPRESSION}[compression]
    ptype = 1 if half else 2
    dt = _PIXEL_TYPES[ptype]
    order = sorted(range(c), key=lambda k: names[k])  # channels are stored in name order
    chl = b"".join(names[k].encode() + b"\0" + struct.pack("<iB3xii", ptype, 0, 1, 1) for k in order) + b"\0"
    box = struct.pack("<4i", 0, 0, w - 1, h - 1)
    header = (struct.pack("<ii", EXR_MAGIC, 2) + _attr("channels", "chlist", chl)
              + _attr("compression", "compression", bytes([comp])) + _attr("dataWindow", "box2i", box)
              + _attr("displayWindow", "box2i", box) + _attr("lineOrder", "lineOrder", b"\0")
              + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
              + _attr("screenWindowCenter", "v2f", struct.pack("<2f", 0.0, 0.0))
              + _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    planes = np.ascontiguousarray(a.astype(dt)[..., order].transpose(0, 2, 1))  # [H, C, W]
    lpc = _LINES_PER_CHUNK[comp]
    chunks = []
    for y0 in range(0, h, lpc):
        raw = planes[y0:y0 + lpc].tobytes()
        if comp == ZIP_COMPRESSION:
            z = zlib.compress(_zip_predict(raw))
            data = z if len(z) < len(raw) else raw
        else:
            data = raw
        chunks.append(struct.pack("<ii", y0, len(data)) + data)
    off = len(header) + 8 * len(chunks)
    table = bytearray()
    for ch in chunks:
        table += struct.pack("<Q", off)
        off += len(ch)
    Path(path).write_bytes(header + bytes(table) + b"".join(chunks))


def _parse_header(b: bytes):
    magic, version = struct.unpack("<ii", b[:8])
    if magic != EXR_MAGIC:
        raise ValueError("not an OpenEXR file")
    if version & 0x200:
        raise ValueError("tiled OpenEXR files are not supported")
    if version & 0x1800:
        raise ValueError("multi-part / deep OpenEXR files are not supported")
    i = 8
    attrs = {}
    while b[i] != 0:
        j = b.index(0, i)
        k = b.index(0, j + 1)
        size = struct.unpack("<i", b[k + 1:k + 5])[0]
        attrs[b[i:j].decode()] = (b[j + 1:k].decode(), b[k + 5:k + 5 + size])
        i = k + 5 + size
    return attrs, i + 1


def read_exr(path) -> tuple[np.ndarray, list[str]]:
    """Reads a scanline OpenEXR file: returns ([H, W, C] float32 (uint32 for UINT channels), channel
    names in file order)."""
    b = Path(path).read_bytes()
    attrs, pos = _parse_header(b)
    chl = attrs["channels"][1]
    chans = []
    i = 0
    while chl[i] != 0:
        j = chl.index(0, i)
        ptype, _, xs, ys = struct.unpack("<iB3xii", chl[j + 1:j + 17])
        if (xs, ys) != (1, 1):
            raise ValueError("subsampled EXR channels are not supported")
        chans.append((chl[i:j].decode(), _PIXEL_TYPES[ptype]))
        i = j + 17
    comp = attrs["compression"][1][0]
    if comp not in _LINES_PER_CHUNK:
        raise ValueError(f"EXR compression {comp} is not supported (none, RLE, ZIPS and ZIP are)")
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lpc = _LINES_PER_CHUNK[comp]
    nchunks = (h + lpc - 1) // lpc
    offsets = struct.unpack(f"<{nchunks}Q", b[pos:pos + 8 * nchunks])
    line_bytes = sum(dt.itemsize for _, dt in chans) * w
    planes = {n: np.empty((h, w), dt) for n, dt in chans}
    for off in offsets:
        y, size = struct.unpack("<ii", b[off:off + 8])
        data = b[off + 8:off + 8 + size]
        nl = min(lpc, y1 + 1 - y)
        expect = nl * line_bytes
        if size < expect:
            data = _rle_decode(data, expect) if comp == RLE_COMPRESSION else _zip_unpredict(zlib.decompress(data))
        if len(data) != expect:
            raise ValueError("corrupt EXR chunk")
        p = 0
        for line in range(nl):
            for n, dt in chans:
                planes[n][y - y0 + line] = np.frombuffer(data, dt, w, p)
                p += dt.itemsize * w
    names = [n for n, _ in chans]
    out_dt = np.uint32 if all(dt == _PIXEL_TYPES[0] for _, dt in chans) else np.float32
    return np.stack([planes[n].astype(out_dt) for n in names], axis=-1), names


def read_hdr(path) -> np.ndarray:
    """Radiance RGBE (.hdr) as [H, W, 3] float32: run-length encoded (new-style) or flat scanlines,
    standard orientation "-Y H +X W"; a texel is byte * 2^(e - 136), zero when e = 0 (the rgbe.c
    conversion FreeImage's HDR loader uses, Bitmap.cpp). Flat scanlines are read as literal
    pixels, as FreeImage's rgbe_ReadPixels reads them (no old-style run markers). Truncated or
    inconsistent data raises ValueError."""
    b = Path(path).read_bytes()
    pos = 0
    first = True

    def line_end(at):
        end = b.find(b"\n", at)
        if end < 0:
            raise ValueError("truncated HDR header")
        return end

    while True:
        end = line_end(pos)
        line = b[pos:end].decode("latin-1").strip()
        pos = end + 1
        if first and not line.startswith("#?"):
            raise ValueError("not a Radiance HDR file")
        first = False
        if line.startswith("FORMAT=") and line != "FORMAT=32-bit_rle_rgbe":
            raise ValueError(f"unsupported HDR format {line}")
        if line == "":
            break
    end = line_end(pos)
    res = b[pos:end].decode("latin-1").split()
    pos = end + 1
    if len(res) != 4 or res[0] != "-Y" or res[2] != "+X":
        raise ValueError(f"unsupported HDR orientation {' '.join(res)}")
    try:
        H, W = int(res[1]), int(res[3])
    except ValueError:
        raise ValueError(f"bad HDR resolution {' '.join(res)}") from None
    if H <= 0 or W <= 0:
        raise ValueError(f"bad HDR resolution {' '.join(res)}")
    n_bytes = len(b)

    def need(k):
        if pos + k > n_bytes:
            raise ValueError("corrupt HDR scanline (truncated data)")

    rgbe = np.zeros((H, W, 4), np.uint8)
    for y in range(H):
        need(4)
        if 8 <= W < 0x8000 and b[pos] == 2 and b[pos + 1] == 2 and ((b[pos + 2] << 8) | b[pos + 3]) == W:
            pos += 4
            for c in range(4):
                x = 0
                while x < W:
                    need(1)
                    n = b[pos]
                    pos += 1
                    if n > 128:
                        n -= 128
                        if x + n > W:
                            raise ValueError("corrupt HDR scanline (run past the row)")
                        need(1)
                        rgbe[y, x:x + n, c] = b[pos]
                        pos += 1
                    else:
                        if n == 0 or x + n > W:
                            raise ValueError("corrupt HDR scanline")
                        need(n)
                        rgbe[y, x:x + n, c] = np.frombuffer(b, np.uint8, n, pos)
                        pos += n
                    x += n
        else:
            need(4 * W)
            rgbe[y] = np.frombuffer(b, np.uint8, 4 * W, pos).reshape(W, 4)
            pos += 4 * W
    e = rgbe[..., 3].astype(np.int32)
    f = np.where(e > 0, np.ldexp(np.float32(1.0), e - 136).astype(np.float32), np.float32(0.0))
    return (rgbe[..., :3].astype(np.float32) * f[..., None]).astype(np.float32)


def read_rgba(path) -> np.ndarray:
    """Any supported file as [H, W, 4] float32 RGBA, missing alpha = 1 (FreeImage_ConvertToRGBAF,
    ImageCompare.cpp:68-110): EXR channels R, G, B, A (Y as grey), PFM, or 8-bit PNG / other
    formats PIL reads, scaled by 1/255."""
    p = Path(path)
    suf = p.suffix.lower()
    if suf == ".exr":
        img, names = read_exr(p)
        out = np.zeros(img.shape[:2] + (4,), np.float32)
        out[..., 3] = 1.0
        idx = {n: k for k, n in enumerate(names)}
        if "Y" in idx and "R" not in idx:
            for c in range(3):
                out[..., c] = img[..., idx["Y"]]
        for c, n in enumerate("RGBA"):
            if n in idx:
                out[..., c] = img[..., idx[n]]
        return out
    if suf == ".hdr":
        img = read_hdr(p)
        out = np.ones(img.shape[:2] + (4,), np.float32)
        out[..., :3] = img
        return out
    if suf == ".pfm":
        img = read_pfm(p)
        out = np.ones(img.shape[:2] + (4,), np.float32)
        out[..., :img.shape[2]] = img
        if img.shape[2] == 1:
            out[..., 1] = out[..., 2] = img[..., 0]
        return out
    from PIL import Image
    with Image.open(p) as im:
        a = np.asarray(im.convert("RGBA"), np.float32) / np.float32(255.0)
    return a


def write_pfm(path, img):
    """Portable float map, RGB (or greyscale), little endian, rows bottom to top."""
    a = np.asarray(img, np.float32)
    if a.ndim == 3 and a.shape[2] == 4:
        a = a[..., :3]
    grey = a.ndim == 2 or a.shape[2] == 1
    a = a.reshape(a.shape[0], a.shape[1], 1 if grey else 3)
    head = f"{'Pf' if grey else 'PF'}\n{a.shape[1]} {a.shape[0]}\n-1.0\n".encode()
    Path(path).write_bytes(head + np.ascontiguousarray(a[::-1]).astype("<f4").tobytes())


def read_pfm(path) -> np.ndarray:
    b = Path(path).read_bytes()
    parts = b.split(b"\n", 3)
    kind, dims, scale = parts[0].strip(), parts[1].split(), float(parts[2])
    c = 3 if kind == b"PF" else 1
    w, h = int(dims[0]), int(dims[1])
    dt = "<f4" if scale < 0 else ">f4"
    a = np.frombuffer(parts[3], dt, w * h * c).reshape(h, w, c)
    return a[::-1].astype(np.float32)


def write_png(path, img, alpha: bool = False):
    """8-bit PNG with channel values int(v * 255) clamped to [0, 255] (ImageCompare.cpp:163-177)."""
    from PIL import Image
    a = np.asarray(img, np.float32)
    c = 4 if alpha else 3
    q = np.clip((a[..., :c] * np.float32(255.0)).astype(np.int64), 0, 255).astype(np.uint8)
    Image.fromarray(np.ascontiguousarray(q), "RGBA" if alpha else "RGB").save(path)


def save_image(path, img, alpha: bool = True, half: bool = False, compression: str = "none"):
    """Saves [H, W, 4] float32 by file extension: .exr, .pfm (RGB), or an 8-bit format (.png ...)."""
    p = Path(path)
    a = np.asarray(img, np.float32)
    suf = p.suffix.lower()
    if suf == ".exr":
        write_exr(p, a if alpha else a[..., :3], half=half, compression=compression)
    elif suf == ".pfm":
        write_pfm(p, a)
    else:
        write_png(p, a, alpha=alpha)
