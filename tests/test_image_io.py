"""Image files (falcor_amd.image_io) and the ImageCompare command line without a GPU.

Fixtures: tests/golden/texture7.exr and texture8.exr are the reference's own EXR test textures
(data/tests/, uncompressed FLOAT RGBA); their contents are pinned by the expectations of
Source/Tools/FalcorTest/Tests/Utils/TextureAnalyzerTests.cpp:100-119. ZIP-compressed files are
checked by round trip only (the reference holds no ZIP-compressed EXR file).
"""
from pathlib import Path

import numpy as np
import pytest

from falcor_amd import image_io as I

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_texture7_known_contents():
    # "R constant, G inf, B NaN, A varying, all channels positive": value (1/8, 2/8, 3/8, 4/8),
    # min (1/8, 2/8, 3/8, 1/16), max (1/8, inf, 3/8, 4/8).
    img = I.read_rgba(GOLDEN / "texture7.exr")
    assert img.shape == (33, 31, 4)
    assert np.array_equal(img[0, 0], [0.125, 0.25, 0.375, 0.5])
    r, g, b, a = (img[..., k] for k in range(4))
    assert np.all(r == 0.125)
    assert np.isinf(g).any() and np.nanmin(g) == 0.25 and np.nanmax(g) == np.inf
    assert np.isnan(b).any() and np.nanmin(b) == 0.375 and np.nanmax(b) == 0.375
    assert a.min() == 0.0625 and a.max() == 0.5
    assert np.all(img[~np.isnan(img)] > 0)


def test_texture8_known_contents():
    # "R varying neg, G constant neg, B varying neg, A varying pos/neg", first texel (-19, -17, -15, -13).
    img = I.read_rgba(GOLDEN / "texture8.exr")
    assert img.shape == (99, 1, 4)
    assert np.array_equal(img[0, 0], [-19, -17, -15, -13])
    r, g, b, a = (img[..., k] for k in range(4))
    assert np.all(r < 0) and len(np.unique(r)) > 1
    assert np.all(g == -17)
    assert np.all(b < 0) and len(np.unique(b)) > 1
    assert a.min() < 0 < a.max()


@pytest.mark.parametrize("name", ["texture7.exr", "texture8.exr"])
def test_uncompressed_float_writer_reproduces_reference_files(tmp_path, name):
    src = GOLDEN / name
    I.write_exr(tmp_path / "out.exr", I.read_rgba(src))
    assert (tmp_path / "out.exr").read_bytes() == src.read_bytes()


@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("compression", ["none", "zip"])
@pytest.mark.parametrize("channels", [1, 3, 4])
def test_exr_round_trip(tmp_path, half, compression, channels):
    rng = np.random.default_rng(channels)
    img = (rng.standard_normal((37, 21, channels)) * 10).astype(np.float32)
    img[0, 0, 0] = np.inf
    img[1, 0, 0] = np.nan
    I.write_exr(tmp_path / "x.exr", img, half=half, compression=compression)
    back, names = I.read_exr(tmp_path / "x.exr")
    want = img.astype(np.float16).astype(np.float32) if half else img
    order = {1: ["Y"], 3: ["R", "G", "B"], 4: ["R", "G", "B", "A"]}[channels]
    back = back[..., [names.index(n) for n in order]]
    assert np.array_equal(back.view(np.uint32), want.view(np.uint32))


def test_zip_compresses_smooth_images(tmp_path):
    y, x = np.mgrid[0:64, 0:64].astype(np.float32)
    img = np.stack([x / 64, y / 64, (x + y) / 128], axis=-1).astype(np.float32)
    I.write_exr(tmp_path / "z.exr", img, compression="zip")
    I.write_exr(tmp_path / "n.exr", img)
    assert (tmp_path / "z.exr").stat().st_size < 0.8 * (tmp_path / "n.exr").stat().st_size
    assert np.array_equal(I.read_rgba(tmp_path / "z.exr")[..., :3], img)


def test_piz_is_refused(tmp_path):
    p = tmp_path / "p.exr"
    I.write_exr(p, np.zeros((2, 2, 3), np.float32))
    b = bytearray(p.read_bytes())
    k = b.index(b"compression\0compression\0") + len(b"compression\0compression\0") + 4
    b[k] = I.PIZ_COMPRESSION
    p.write_bytes(bytes(b))
    with pytest.raises(ValueError, match="compression 4"):
        I.read_exr(p)


def test_pfm_and_png(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.random((5, 7, 4)).astype(np.float32)
    I.save_image(tmp_path / "a.pfm", img)
    back = I.read_rgba(tmp_path / "a.pfm")
    assert np.array_equal(back[..., :3], img[..., :3]) and np.all(back[..., 3] == 1)
    I.save_image(tmp_path / "a.png", img, alpha=True)
    q = I.read_rgba(tmp_path / "a.png")
    # int(v * 255) truncates (ImageCompare.cpp:169-172).
    assert np.array_equal(np.round(q * 255).astype(int), np.clip((img * np.float32(255)).astype(int), 0, 255))


def test_image_compare_cli_without_device(capsys):
    from falcor_amd.image_compare import main
    assert main(["-l"]) == 0
    out = capsys.readouterr().out
    assert "mse - Mean Squared Error" in out and "mape" in out
    assert main(["-m", "nope", "a.exr", "b.exr"]) == 1
    assert main([]) == 1


def _rgbe_encode(rgbe, rle=True):
    """Radiance scanlines of [H, W, 4] uint8 quadruples: new-style run-length or flat."""
    H, W, _ = rgbe.shape
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {H} +X {W}\n".encode())
    for y in range(H):
        if not rle:
            out += rgbe[y].tobytes()
            continue
        out += bytes([2, 2, W >> 8, W & 255])
        for c in range(4):
            ch = rgbe[y, :, c]
            x = 0
            while x < W:
                r = 1
                while x + r < W and r < 127 and ch[x + r] == ch[x]:
                    r += 1
                if r >= 3:
                    out += bytes([128 + r, ch[x]])
                    x += r
                else:
                    n = min(128, W - x)
                    out += bytes([n]) + ch[x:x + n].tobytes()
                    x += n
    return bytes(out)


@pytest.mark.parametrize("rle", [True, False])
def test_hdr_reader_decodes_rgbe(tmp_path, rle):
    from falcor_amd import image_io as I
    rng = np.random.default_rng(1)
    H, W = 5, 37
    rgbe = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    rgbe[:, 10:20] = rgbe[:, 10:11]           # runs
    rgbe[2, :, 3] = 0                         # zero exponent: black
    rgbe[3, 4:6, :3] = 1                      # (1, 1, 1, e) texels are literal pixels, as FreeImage reads them
    (tmp_path / "a.hdr").write_bytes(_rgbe_encode(rgbe, rle))
    got = I.read_hdr(tmp_path / "a.hdr")
    e = rgbe[..., 3].astype(np.int64)
    want = np.where(e[..., None] > 0, rgbe[..., :3] * np.ldexp(1.0, e - 136)[..., None], 0.0)
    np.testing.assert_array_equal(got, want.astype(np.float32))
    np.testing.assert_array_equal(I.read_rgba(tmp_path / "a.hdr")[..., :3], got)


@pytest.mark.parametrize("rle", [True, False])
def test_hdr_reader_rejects_truncated_files(tmp_path, rle):
    from falcor_amd import image_io as I
    rng = np.random.default_rng(2)
    rgbe = rng.integers(1, 256, (4, 16, 4), dtype=np.uint8)
    data = _rgbe_encode(rgbe, rle)
    header = data.index(b"+X 16\n") + len(b"+X 16\n")
    for cut in (header + 1, header + 3, header + 17, len(data) - 1):
        (tmp_path / "t.hdr").write_bytes(data[:cut])
        with pytest.raises(ValueError, match="corrupt HDR scanline"):
            I.read_hdr(tmp_path / "t.hdr")
    (tmp_path / "h.hdr").write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n")
    with pytest.raises(ValueError, match="truncated HDR header"):
        I.read_hdr(tmp_path / "h.hdr")


def test_reference_default_hdr_env_map():
    from pathlib import Path
    from falcor_amd import image_io as I
    p = Path("/root/reference/Scene/EnvMap/Default.hdr")
    if not p.exists():
        pytest.skip("reference env map not present")
    img = I.read_hdr(p)
    assert img.shape == (512, 1024, 3) and np.isfinite(img).all() and (img >= 0).all()
    # Every nonzero texel has a normalised mantissa: its largest channel is in [128, 256) * 2^(e - 136).
    m = img.max(-1)
    nz = m > 0
    e = np.floor(np.log2(m[nz])) + 1
    frac = m[nz] / np.exp2(e)
    assert ((frac >= 0.5) & (frac < 1.0)).all()
