"""OpenEXR scanline files without the OpenEXR library (the reference uses it through
scripts/exr.py, which is not available here).

Reads and writes single-part scanline images with NO_COMPRESSION, RLE_COMPRESSION,
ZIPS_COMPRESSION (one scanline per chunk) or ZIP_COMPRESSION (16 scanlines per chunk), pixel types
HALF, FLOAT and UINT, following the published file layout: magic 20000630, version 2, a header of
(name, type, size, value) attributes ended by a zero byte, one 64-bit offset per chunk, then chunks
of (first scanline, byte count, data) whose uncompressed data holds, per scanline, every channel in
name order. ZIP and RLE first split the bytes into even and odd positions and replace each byte by
its difference to the previous one (+128); ZIP then deflates (zlib), RLE writes runs as a signed
count byte (n >= 0: n + 1 copies of the next byte; n < 0: -n literal bytes). A chunk whose
compressed form is not smaller is stored raw. PIZ, PXR24, B44 and DWA files, tiled, deep and
multi-part files are rejected with an error.

The interface follows scripts/exr.py: ``write(path, data, channel_names=None, precision=FLOAT,
compression=...)`` with the same default channel names (1: Z, 2: X Y, 3: R G B, 4: R G B A) and
"group.channel" names for dict input; ``read_all(path)`` returns {group: [H, W, C] array} with
"default" for the channels without a group, each group's channels in the reference's order
(R, G, B, A, then X, Y, Z, then other names -- sorted by name, where the reference's mixed int/str
sort key would raise). UINT reads as uint32 (the reference maps it to uint8).
"""
from __future__ import annotations

import struct
import zlib
from collections import OrderedDict

import numpy as np

MAGIC = 20000630

UINT, HALF, FLOAT = 0, 1, 2
_PIXEL_DTYPE = {UINT: np.dtype("<u4"), HALF: np.dtype("<f2"), FLOAT: np.dtype("<f4")}
_PRECISION_NAMES = {"UINT": UINT, "HALF": HALF, "FLOAT": FLOAT}

NO_COMPRESSION, RLE_COMPRESSION, ZIPS_COMPRESSION, ZIP_COMPRESSION = 0, 1, 2, 3
PIZ_COMPRESSION, PXR24_COMPRESSION = 4, 5
_LINES_PER_CHUNK = {NO_COMPRESSION: 1, RLE_COMPRESSION: 1, ZIPS_COMPRESSION: 1, ZIP_COMPRESSION: 16}
_COMPRESSION_NAMES = {"NO_COMPRESSION": 0, "RLE_COMPRESSION": 1, "ZIPS_COMPRESSION": 2, "ZIP_COMPRESSION": 3,
                      "PIZ_COMPRESSION": 4, "PXR24_COMPRESSION": 5, "B44_COMPRESSION": 6, "B44A_COMPRESSION": 7,
                      "DWAA_COMPRESSION": 8, "DWAB_COMPRESSION": 9}

DEFAULT_CHANNEL_NAMES = {1: ["Z"], 2: ["X", "Y"], 3: ["R", "G", "B"], 4: ["R", "G", "B", "A"]}
_RANK = {"r": 10, "g": 20, "b": 30, "a": 40, "x": 110, "y": 120, "z": 130}


class ExrError(ValueError):
    pass


def _as_code(value, table, what):
    if isinstance(value, str):
        if value not in table:
            raise ExrError(f"unknown {what} {value!r}")
        return table[value]
    return int(value)


def channel_sort_key(name: str):
    """Order of scripts/exr.py:_channel_sort_key (R, G, B, A, X, Y, Z first) made total: names
    outside that set follow, sorted by name."""
    return [(0, _RANK[p.lower()], "") if p.lower() in _RANK else (1, 0, p) for p in name.split(".")]


# ---------------------------------------------------------------------------------------------
# byte transforms of ZIP / RLE
# ---------------------------------------------------------------------------------------------
def _reorder_predict(raw: bytes) -> bytes:
    b = np.frombuffer(raw, dtype=np.uint8)
    t = np.concatenate([b[0::2], b[1::2]])
    if t.size > 1:
        d = (t[1:].astype(np.int16) - t[:-1].astype(np.int16) + 128) & 0xFF
        t = np.concatenate([t[:1], d.astype(np.uint8)])
    return t.tobytes()


def _unpredict_unorder(data: bytes) -> bytes:
    t = np.frombuffer(data, dtype=np.uint8).astype(np.int64)
    if t.size > 1:
        t[1:] -= 128
        t = np.cumsum(t) & 0xFF
    t = t.astype(np.uint8)
    half = (t.size + 1) // 2
    out = np.empty_like(t)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def _rle_compress(data: bytes) -> bytes:
    out = bytearray()
    n, i = len(data), 0
    while i < n:
        j = i + 1
        while j < n and data[j] == data[i] and j - i < 128:
            j += 1
        if j - i >= 3:  # a run
            out.append(j - i - 1)
            out.append(data[i])
            i = j
            continue
        j = i  # literal bytes up to the next run of 3 or 127 bytes
        while j < n and j - i < 127:
            if j + 2 < n and data[j] == data[j + 1] == data[j + 2]:
                break
            j += 1
        out.append((-(j - i)) & 0xFF)
        out += data[i:j]
        i = j
    return bytes(out)


def _rle_uncompress(data: bytes, size: int) -> bytes:
    out = bytearray()
    i = 0
    while i < len(data):
        c = struct.unpack_from("b", data, i)[0]
        i += 1
        if c < 0:
            out += data[i:i - c]
            i -= c
        else:
            out += bytes([data[i]]) * (c + 1)
            i += 1
    if len(out) != size:
        raise ExrError("RLE chunk decodes to the wrong size")
    return bytes(out)


def _compress(raw: bytes, comp: int) -> bytes:
    if comp == NO_COMPRESSION:
        return raw
    t = _reorder_predict(raw)
    packed = zlib.compress(t) if comp in (ZIP_COMPRESSION, ZIPS_COMPRESSION) else _rle_compress(t)
    return packed if len(packed) < len(raw) else raw


def _uncompress(data: bytes, comp: int, size: int) -> bytes:
    if comp == NO_COMPRESSION or len(data) == size:
        return data
    if comp in (ZIP_COMPRESSION, ZIPS_COMPRESSION):
        t = zlib.decompress(data)
        if len(t) != size:
            raise ExrError("ZIP chunk decodes to the wrong size")
    else:
        t = _rle_uncompress(data, size)
    return _unpredict_unorder(t)


# ---------------------------------------------------------------------------------------------
# header
# ---------------------------------------------------------------------------------------------
def _attr(name: str, type_: str, value: bytes) -> bytes:
    return name.encode() + b"\0" + type_.encode() + b"\0" + struct.pack("<i", len(value)) + value


def _header(width: int, height: int, channels, comp: int) -> bytes:
    chl = b"".join(c.encode() + b"\0" + struct.pack("<iB3xii", t, 0, 1, 1) for c, t in channels) + b"\0"
    box = struct.pack("<iiii", 0, 0, width - 1, height - 1)
    return (_attr("channels", "chlist", chl) + _attr("compression", "compression", bytes([comp]))
            + _attr("dataWindow", "box2i", box) + _attr("displayWindow", "box2i", box)
            + _attr("lineOrder", "lineOrder", b"\0") + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
            + _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
            + _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")


def _parse_header(buf: bytes):
    if len(buf) < 8:
        raise ExrError("not an EXR file (too short)")
    magic, version = struct.unpack_from("<iI", buf, 0)
    if magic != MAGIC:
        raise ExrError("not an EXR file (bad magic number)")
    if version & 0xFF != 2:
        raise ExrError(f"unsupported EXR version {version & 0xFF}")
    if version & 0x1A00:  # tiled (0x200), deep (0x800), multi-part (0x1000)
        raise ExrError("only single-part scanline EXR files are supported")
    pos, attrs = 8, {}
    while True:
        end = buf.index(b"\0", pos)
        name = buf[pos:end].decode()
        pos = end + 1
        if not name:
            break
        end = buf.index(b"\0", pos)
        type_ = buf[pos:end].decode()
        size = struct.unpack_from("<i", buf, end + 1)[0]
        attrs[name] = (type_, buf[end + 5:end + 5 + size])
        pos = end + 5 + size
    for need in ("channels", "compression", "dataWindow"):
        if need not in attrs:
            raise ExrError(f"EXR header lacks the {need!r} attribute")
    chl, channels, p = attrs["channels"][1], [], 0
    while chl[p] != 0:
        end = chl.index(b"\0", p)
        name = chl[p:end].decode()
        t, _, xs, ys = struct.unpack_from("<iB3xii", chl, end + 1)
        if (xs, ys) != (1, 1):
            raise ExrError("subsampled channels are not supported")
        if t not in _PIXEL_DTYPE:
            raise ExrError(f"unknown pixel type {t}")
        channels.append((name, t))
        p = end + 17
    comp = attrs["compression"][1][0]
    if comp not in _LINES_PER_CHUNK:
        names = {v: k for k, v in _COMPRESSION_NAMES.items()}
        raise ExrError(f"{names.get(comp, comp)} is not supported (NO, RLE, ZIPS and ZIP are)")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    return pos, channels, comp, (x0, y0, x1, y1)


# ---------------------------------------------------------------------------------------------
# files
# ---------------------------------------------------------------------------------------------
def write_channels(path: str, channels, compression=ZIP_COMPRESSION) -> None:
    """channels: {name: (2-D array [H, W], pixel type)}; written in name order."""
    comp = _as_code(compression, _COMPRESSION_NAMES, "compression")
    if comp not in _LINES_PER_CHUNK:
        raise ExrError("only NO, RLE, ZIPS and ZIP compression are written")
    items = sorted(channels.items())
    if not items:
        raise ExrError("no channels to write")
    height, width = np.asarray(items[0][1][0]).shape
    planes = []
    for name, (arr, t) in items:
        a = np.asarray(arr)
        if a.shape != (height, width):
            raise ExrError(f"channel {name!r} has shape {a.shape}, expected {(height, width)}")
        planes.append(np.ascontiguousarray(a.astype(_PIXEL_DTYPE[t])))
    lines = _LINES_PER_CHUNK[comp]
    header = struct.pack("<iI", MAGIC, 2) + _header(width, height, [(n, t) for n, (_, t) in items], comp)
    chunks = []
    for y in range(0, height, lines):
        raw = b"".join(b"".join(p[r].tobytes() for p in planes) for r in range(y, min(y + lines, height)))
        data = _compress(raw, comp)
        chunks.append(struct.pack("<ii", y, len(data)) + data)
    offsets, pos = [], len(header) + 8 * len(chunks)
    for c in chunks:
        offsets.append(pos)
        pos += len(c)
    with open(path, "wb") as f:
        f.write(header)
        f.write(struct.pack(f"<{len(offsets)}Q", *offsets))
        for c in chunks:
            f.write(c)


def read_channels(path: str) -> "OrderedDict[str, np.ndarray]":
    """{name: [H, W] array} of every channel, in the file's (name) order."""
    with open(path, "rb") as f:
        buf = f.read()
    pos, channels, comp, (x0, y0, x1, y1) = _parse_header(buf)
    width, height = x1 - x0 + 1, y1 - y0 + 1
    lines = _LINES_PER_CHUNK[comp]
    nchunks = (height + lines - 1) // lines
    offsets = struct.unpack_from(f"<{nchunks}Q", buf, pos)
    out = OrderedDict((n, np.empty((height, width), dtype=_PIXEL_DTYPE[t])) for n, t in channels)
    row_bytes = [width * _PIXEL_DTYPE[t].itemsize for _, t in channels]
    for off in offsets:
        y, size = struct.unpack_from("<ii", buf, off)
        r0 = y - y0
        nrows = min(lines, height - r0)
        if r0 < 0 or nrows <= 0:
            raise ExrError("chunk outside the data window")
        raw = _uncompress(buf[off + 8:off + 8 + size], comp, nrows * sum(row_bytes))
        p = 0
        for r in range(r0, r0 + nrows):
            for (name, t), nb in zip(channels, row_bytes):
                out[name][r] = np.frombuffer(raw, dtype=_PIXEL_DTYPE[t], count=width, offset=p)
                p += nb
    return out


def write(path: str, data, channel_names=None, precision="FLOAT", compression=ZIP_COMPRESSION) -> None:
    """scripts/exr.py:write: ``data`` an [H, W] or [H, W, C] array (channel names given or the
    defaults for C <= 4) or a dict {group: array} ("default" for ungrouped channels)."""
    def as3(m):
        m = np.asarray(m)
        if m.ndim == 2:
            m = m[:, :, None]
        if m.ndim != 3:
            raise ExrError("Invalid number of dimensions for the `matrix` argument.")
        return m

    def names_for(given, depth):
        if given:
            if len(given) != depth:
                raise ExrError(f"The provided channel names have the wrong length ({depth} vs {len(given)}).")
            return list(given)
        if depth in DEFAULT_CHANNEL_NAMES:
            return DEFAULT_CHANNEL_NAMES[depth]
        raise ExrError(f"There are no suitable default channel names for data of depth {depth}")

    groups = data if isinstance(data, dict) else {"default": data}
    given = channel_names if isinstance(channel_names, dict) else {"default": channel_names}
    prec = precision if isinstance(precision, dict) else {g: precision for g in groups}
    channels = {}
    for g, m in groups.items():
        m = as3(m)
        t = _as_code(prec.get(g, "FLOAT"), _PRECISION_NAMES, "precision")
        for i, c in enumerate(names_for(given.get(g), m.shape[2])):
            channels[c if g == "default" else f"{g}.{c}"] = (m[:, :, i], t)
    write_channels(path, channels, compression)


def read_all(path: str) -> dict:
    """scripts/exr.py:read_all: {group: [H, W, C] array}, channels of a group in the reference's order."""
    ch = read_channels(path)
    groups = {}
    for name in sorted(ch, key=channel_sort_key):
        g = name.split(".")[0] if "." in name else "default"
        groups.setdefault(g, []).append(ch[name])
    return {g: np.stack(v, axis=-1) for g, v in groups.items()}


def read(path: str, group: str = "default") -> np.ndarray:
    """The [H, W, C] array of one group (scripts/exr.py:read with a group name)."""
    groups = read_all(path)
    if group not in groups:
        raise ExrError(f"no channels in group {group!r} (have {sorted(groups)})")
    return groups[group]
