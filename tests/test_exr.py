"""xdn.exr (OpenEXR scanline files without the OpenEXR library), the EXR scene loader and the
driver's image output. The OpenEXR library is absent here, so the module is pinned to the published
file layout by a byte-level fixture laid out by hand, to hand-computed known answers of the ZIP/RLE
byte transforms, and by round trips; reading files written by the OpenEXR library itself stays
parity unpinned."""
import os
import struct

import numpy as np
import pytest
import torch

from xdn import data, driver, exr


def test_uncompressed_file_layout_byte_for_byte(tmp_path):
    """A 2x2 one-channel FLOAT image, no compression: magic, version, the eight header attributes,
    the offset table (one chunk per scanline) and the two chunks, each field written out by hand."""
    img = np.array([[1.0, 2.0], [3.0, -0.5]], dtype=np.float32)
    p = str(tmp_path / "z.exr")
    exr.write(p, img, compression=exr.NO_COMPRESSION)
    hdr = b""
    hdr += b"\x76\x2f\x31\x01" + b"\x02\x00\x00\x00"
    hdr += b"channels\x00chlist\x00" + (19).to_bytes(4, "little")  # one 18-byte entry + terminator
    hdr += b"Z\x00" + (2).to_bytes(4, "little") + b"\x00\x00\x00\x00" + (1).to_bytes(4, "little") * 2 + b"\x00"
    hdr += b"compression\x00compression\x00" + (1).to_bytes(4, "little") + b"\x00"
    win = (0).to_bytes(4, "little") * 2 + (1).to_bytes(4, "little") * 2
    hdr += b"dataWindow\x00box2i\x00" + (16).to_bytes(4, "little") + win
    hdr += b"displayWindow\x00box2i\x00" + (16).to_bytes(4, "little") + win
    hdr += b"lineOrder\x00lineOrder\x00" + (1).to_bytes(4, "little") + b"\x00"
    hdr += b"pixelAspectRatio\x00float\x00" + (4).to_bytes(4, "little") + b"\x00\x00\x80\x3f"
    hdr += b"screenWindowCenter\x00v2f\x00" + (8).to_bytes(4, "little") + b"\x00" * 8
    hdr += b"screenWindowWidth\x00float\x00" + (4).to_bytes(4, "little") + b"\x00\x00\x80\x3f"
    hdr += b"\x00"
    chunk0 = (0).to_bytes(4, "little") + (8).to_bytes(4, "little") + b"\x00\x00\x80\x3f" + b"\x00\x00\x00\x40"
    chunk1 = (1).to_bytes(4, "little") + (8).to_bytes(4, "little") + b"\x00\x00\x40\x40" + b"\x00\x00\x00\xbf"
    off0 = len(hdr) + 16
    want = hdr + off0.to_bytes(8, "little") + (off0 + len(chunk0)).to_bytes(8, "little") + chunk0 + chunk1
    got = open(p, "rb").read()
    assert got == want
    np.testing.assert_array_equal(exr.read(p)[:, :, 0], img)


def test_zip_byte_transform_known_answer():
    """Even bytes then odd bytes, then each byte minus its predecessor plus 128 (mod 256)."""
    raw = bytes([0x10, 0x20, 0x30, 0x40, 0x50])
    t = exr._reorder_predict(raw)
    assert t == bytes([0x10, 0xA0, 0xA0, 0x50, 0xA0])
    assert exr._unpredict_unorder(t) == raw
    rng = np.random.default_rng(3)
    for n in (0, 1, 2, 7, 64, 1001):
        r = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert exr._unpredict_unorder(exr._reorder_predict(r)) == r


def test_rle_known_answer():
    """A run of n >= 3 equal bytes is (n - 1, byte); literal bytes are (-count, bytes...)."""
    assert exr._rle_compress(b"aaaaab") == b"\x04a\xffb"
    assert exr._rle_compress(b"abc") == b"\xfdabc"
    assert exr._rle_uncompress(b"\x04a\xffb", 6) == b"aaaaab"
    assert exr._rle_compress(b"x" * 300) == b"\x7fx\x7fx\x2bx"
    rng = np.random.default_rng(5)
    for n in (1, 3, 130, 1000):
        r = bytes(rng.choice([0, 0, 0, 7, 9], n).astype(np.uint8))
        assert exr._rle_uncompress(exr._rle_compress(r), n) == r


@pytest.mark.parametrize("comp", ["NO_COMPRESSION", "RLE_COMPRESSION", "ZIPS_COMPRESSION", "ZIP_COMPRESSION"])
@pytest.mark.parametrize("shape", [(1, 1, 3), (17, 5, 1), (37, 23, 4), (33, 64, 3), (40, 7, 6)])
def test_round_trip(tmp_path, comp, shape):
    rng = np.random.default_rng([len(comp), *shape])
    img = (rng.standard_normal(shape) * 10).astype(np.float32)
    img[: shape[0] // 2] = 1.5  # compressible rows beside incompressible ones
    names = None if shape[2] <= 4 else [f"ch{i}" for i in range(shape[2])]
    p = str(tmp_path / "a.exr")
    exr.write(p, img, channel_names=names, compression=comp)
    back = exr.read(p)
    if names:  # the "default" group in the reference's order: names outside R,G,B,A,X,Y,Z by name
        order = sorted(range(shape[2]), key=lambda i: names[i])
        img = img[:, :, order]
    np.testing.assert_array_equal(back, img)


def test_half_uint_and_groups(tmp_path):
    rng = np.random.default_rng(9)
    h = (rng.standard_normal((9, 13)) * 4).astype(np.float16)
    u = rng.integers(0, 2**32, (9, 13), dtype=np.uint64).astype(np.uint32)
    p = str(tmp_path / "m.exr")
    exr.write_channels(p, {"A": (h, exr.HALF), "depth.Z": (u, exr.UINT), "B": (h * 2, exr.HALF),
                           "R": (h + 1, exr.HALF), "G": (h - 1, exr.HALF)}, compression="ZIP_COMPRESSION")
    ch = exr.read_channels(p)
    assert list(ch) == ["A", "B", "G", "R", "depth.Z"]  # file order is by name
    assert ch["A"].dtype == np.float16 and ch["depth.Z"].dtype == np.uint32
    np.testing.assert_array_equal(ch["depth.Z"], u)
    groups = exr.read_all(p)
    assert sorted(groups) == ["default", "depth"]
    np.testing.assert_array_equal(groups["default"], np.stack([h + 1, h - 1, h * 2, h], -1))  # R, G, B, A
    # dict input and precision per group, as scripts/exr.py:write takes them
    q = str(tmp_path / "d.exr")
    exr.write(q, {"default": np.ones((4, 5, 3), np.float32), "normal": np.zeros((4, 5, 2), np.float32)},
              precision={"normal": "HALF"})
    names = list(exr.read_channels(q))
    assert names == ["B", "G", "R", "normal.X", "normal.Y"]
    assert exr.read(q, "normal").dtype == np.float16


def test_rejected_files_and_arguments(tmp_path):
    p = str(tmp_path / "x.exr")
    with pytest.raises(exr.ExrError):
        exr.write(p, np.zeros((2, 2, 3), np.float32), compression="PIZ_COMPRESSION")
    with pytest.raises(exr.ExrError):
        exr.write(p, np.zeros((2, 2, 5), np.float32))  # no default names for 5 channels
    exr.write(p, np.zeros((2, 2, 3), np.float32), compression="NO_COMPRESSION")
    buf = bytearray(open(p, "rb").read())
    bad = str(tmp_path / "bad.exr")
    open(bad, "wb").write(b"\x00" * 4 + bytes(buf[4:]))
    with pytest.raises(exr.ExrError, match="magic"):
        exr.read(bad)
    buf[5] |= 0x02  # tiled flag (bit 9 of the version field)
    open(bad, "wb").write(bytes(buf))
    with pytest.raises(exr.ExrError, match="scanline"):
        exr.read(bad)
    buf[5] &= ~0x02
    i = bytes(buf).index(b"compression\x00compression\x00") + len("compression\x00compression\x00") + 4
    buf[i] = 4  # PIZ
    open(bad, "wb").write(bytes(buf))
    with pytest.raises(exr.ExrError, match="PIZ"):
        exr.read(bad)


def _frames(n, H=12, W=20, seed=0):
    g = torch.Generator().manual_seed(seed)
    out = []
    for _ in range(n):
        f = {t: torch.rand(1, H, W, 3, generator=g) * 4 for t in data.READ_TYPES if t != "opacity"}
        f["opacity"] = torch.ones(1, H, W, 3)
        out.append(f)
    return out


def test_exr_scene_equals_npy_scene(tmp_path):
    frames = _frames(3)
    data.save_exr_scene(str(tmp_path / "exr"), frames, first=4)
    data.save_npy_scene(str(tmp_path / "npy"), frames, first=4)
    a = data.ExrScene(str(tmp_path / "exr"), device="cpu")
    b = data.NpyScene(str(tmp_path / "npy"), device="cpu")
    assert a.frame_ids == b.frame_ids == [4, 5, 6]
    for (fa, xa), (fb, xb) in zip(a, b):
        assert fa == fb and sorted(xa) == sorted(xb)
        for t in xa:
            assert torch.equal(xa[t], xb[t]), t
    assert a[4][0] == 5  # index past the end wraps around
    # an .exr scene is not an .npy scene
    with pytest.raises(FileNotFoundError):
        data.NpyScene(str(tmp_path / "exr"), device="cpu")


def test_nan_inf_cleaning_and_conversion(tmp_path):
    img = np.ones((3, 4, 3), np.float32)
    img[0, 0, 0], img[1, 1, 1], img[2, 2, 2] = np.nan, np.inf, -np.inf
    d = tmp_path / "s"
    d.mkdir()
    for t in data.READ_TYPES:
        exr.write(str(d / f"{t}_0000.exr"), img)
    sc = data.ExrScene(str(d), device="cpu")
    x = sc[0][1]["color"]
    assert x.shape == (1, 3, 4, 3) and torch.isfinite(x).all()
    assert float(x[0, 0, 0, 0]) == 0 and float(x[0, 1, 1, 1]) == 0 and float(x[0, 2, 2, 2]) == 0
    assert data.convert_exr_scene(str(d)) == len(data.READ_TYPES)
    assert data.convert_exr_scene(str(d)) == 0  # existing .npy files are kept
    npy = np.load(str(d / "color_0000.npy"))
    assert npy.shape == (1, 3, 4, 3) and np.array_equal(npy, x.numpy())


def test_driver_image_output(tmp_path):
    rng = np.random.default_rng(2)
    img = rng.standard_normal((6, 5, 3)).astype(np.float32)
    n = driver.write_images(str(tmp_path), [("out", 7, img), ("path", 7, img), ("olsavg", 12, img)])
    assert n == 3 and sorted(os.listdir(tmp_path)) == ["olsavg_0012.exr", "out_0007.exr", "path_0007.exr"]
    np.testing.assert_array_equal(exr.read(str(tmp_path / "out_0007.exr")), np.maximum(img, 0))
    np.testing.assert_array_equal(exr.read(str(tmp_path / "olsavg_0012.exr")), np.maximum(img, 0))
    np.testing.assert_array_equal(exr.read(str(tmp_path / "path_0007.exr")), img)  # not clamped
    assert struct.unpack_from("<i", open(tmp_path / "path_0007.exr", "rb").read(), 0)[0] == exr.MAGIC
    # the composites of main.py:350-359
    f = _frames(1, 4, 6, seed=1)[0]
    res = {"ols1": torch.log1p(f["color"]), "ols2": torch.log1p(f["color2"]), "shown": f["albedo"]}
    ims = driver.frame_images(f, res)
    e, l = f["emissive"][0].numpy(), f["envLight"][0].numpy()
    np.testing.assert_allclose(ims["olsavg"], (f["color"][0].numpy() + f["color2"][0].numpy()) * 0.5 + e + l,
                               rtol=1e-6)
    np.testing.assert_allclose(ims["ref"], (f["ref"] + f["ref_emissive"] + f["ref_envLight"])[0].numpy(), rtol=1e-6)
    assert np.array_equal(ims["out"], f["albedo"][0].numpy())
