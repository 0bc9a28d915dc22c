"""Bitmap textures and the environment-map emitter: image loading
(util/texture.cpp:87-174), bitmap sampling (cuda/texture.h:52-54 with the
texture-object settings of cuda/texture.cpp:60-100), the XML texture and
envmap loaders (resource/scene.cpp:141-222) and EnvMapEmitter
(render/emitter/env.h:6-65, world/emitter.cpp:107-149).

The image files here are written by small encoders in this file, from the
OpenEXR, Radiance RGBE and PNG specifications.  Env-map rendering is pinned
by the irradiance a diffuse floor receives from the map, integrated in
float64 over the bilinearly filtered radiance; the reference's own texel-
corner sampler (MIS_REFERENCE) is a discrete quadrature, so only the exact
sampler (MIS_CONSISTENT) is held to the closed form.
"""
import os
import struct
import zlib

import numpy as np
import pytest

import oracle
import pupiloptixlab_amd as pt


# ---------------------------------------------------------------- encoders
ADAM7 = ((0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2))


def png_bytes(pixels, color_type, depth=8, palette=None, filters=(0, 1, 2, 3, 4), interlace=False):
    """pixels: (h, w, c) uint array in the PNG sample layout.  With
    interlace, the seven Adam7 sub-images are filtered and stored in turn
    (empty ones are left out)."""
    h, w = pixels.shape[:2]
    dt = ">u2" if depth == 16 else "u1"
    bpp = max(1, pixels.shape[2] * depth // 8) if pixels.ndim == 3 else 1

    def filtered(pix):
        raw = bytearray()
        ph, pw = pix.shape[:2]
        prev = bytes((pw * bpp) if depth >= 8 else ((pw * depth + 7) // 8))
        for y in range(ph):
            if depth >= 8:
                row = pix[y].astype(dt).tobytes()
            else:  # packed gray / palette indices
                bits = "".join(format(int(v), f"0{depth}b") for v in pix[y].reshape(-1))
                bits += "0" * (-len(bits) % 8)
                row = bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))
            ft = filters[y % len(filters)]
            out = bytearray()
            for i, v in enumerate(row):
                a = row[i - bpp] if i >= bpp else 0
                b = prev[i]
                c = prev[i - bpp] if i >= bpp else 0
                if ft == 0:
                    pred = 0
                elif ft == 1:
                    pred = a
                elif ft == 2:
                    pred = b
                elif ft == 3:
                    pred = (a + b) // 2
                else:
                    p = a + b - c
                    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                    pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                out.append((v - pred) & 0xFF)
            raw += bytes([ft]) + out
            prev = row
        return raw

    if interlace:
        raw = bytearray()
        for x0, y0, dx, dy in ADAM7:
            sub = pixels[y0::dy, x0::dx]
            if sub.shape[0] and sub.shape[1]:
                raw += filtered(sub)
    else:
        raw = filtered(pixels)

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color_type, 0, 0,
                                                                 1 if interlace else 0))
    if palette is not None:
        data += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    return data + chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")


def rgbe_bytes(rgbe, rle):
    h, w = rgbe.shape[:2]
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        if not rle:
            out += rgbe[y].astype(np.uint8).tobytes()
            continue
        out += bytes([2, 2, w >> 8, w & 0xFF])
        for c in range(4):
            vals = rgbe[y, :, c].astype(np.uint8)
            x = 0
            while x < w:  # runs of equal values become run packets, the rest literals
                r = 1
                while x + r < w and r < 127 and vals[x + r] == vals[x]:
                    r += 1
                if r >= 3:
                    out += bytes([128 + r, vals[x]])
                    x += r
                else:
                    n = min(128, w - x)
                    lit = []
                    while len(lit) < n:
                        j = x + len(lit)
                        if j + 2 < w and vals[j] == vals[j + 1] == vals[j + 2]:
                            break
                        lit.append(vals[j])
                    if not lit:
                        lit = [vals[x]]
                    out += bytes([len(lit)] + [int(v) for v in lit])
                    x += len(lit)
    return bytes(out)


def float_to_half_bits(a):
    return np.asarray(a, np.float16).view(np.uint16)


def exr_bytes(img, compression, half):
    """Scanline EXR with channels A?, B, G, R (alphabetical), compression
    0 NONE, 1 RLE, 2 ZIPS, 3 ZIP; row 0 of img is the first scanline."""
    h, w, nc = img.shape
    names = ["B", "G", "R"] + (["A"] if nc == 4 else [])
    chans = sorted(names)
    src = {"R": 0, "G": 1, "B": 2, "A": 3}
    ptype = 1 if half else 2

    def attr(name, typ, payload):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(payload)) + payload

    chl = b"".join(c.encode() + b"\0" + struct.pack("<iB3xii", ptype, 0, 1, 1) for c in chans) + b"\0"
    hdr = struct.pack("<ii", 20000630, 2)
    hdr += attr("channels", "chlist", chl)
    hdr += attr("compression", "compression", bytes([compression]))
    hdr += attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    hdr += attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    hdr += attr("lineOrder", "lineOrder", b"\0")
    hdr += attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    hdr += attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0))
    hdr += attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0"
    lines = 16 if compression == 3 else 1
    blocks = []
    for y0 in range(0, h, lines):
        raw = bytearray()
        for y in range(y0, min(h, y0 + lines)):
            for c in chans:
                v = img[y, :, src[c]]
                raw += float_to_half_bits(v).tobytes() if half else v.astype("<f4").tobytes()
        data = bytes(raw)
        if compression in (1, 2, 3):
            b = np.frombuffer(data, np.uint8)
            inter = np.concatenate([b[0::2], b[1::2]])  # reorder: even bytes, then odd
            d = inter.astype(np.int16)
            pred = np.empty_like(d)
            pred[0] = d[0]
            pred[1:] = (d[1:] - d[:-1] + 128) & 0xFF
            pb = pred.astype(np.uint8).tobytes()
            if compression == 1:
                enc = bytearray()
                i = 0
                while i < len(pb):  # literal runs only (valid RLE, not minimal)
                    n = min(127, len(pb) - i)
                    enc += bytes([(256 - n) & 0xFF]) + pb[i:i + n]
                    i += n
                comp = bytes(enc)
            else:
                comp = zlib.compress(pb)
            if len(comp) < len(data):
                data = comp
        blocks.append(struct.pack("<ii", y0, len(data)) + data)
    table_start = len(hdr)
    offs, pos = [], table_start + 8 * len(blocks)
    for b in blocks:
        offs.append(pos)
        pos += len(b)
    return hdr + struct.pack(f"<{len(offs)}Q", *offs) + b"".join(blocks)


def write(tmp_path, name, data):
    p = tmp_path / name
    if isinstance(data, str):
        p.write_text(data)
    else:
        p.write_bytes(data)
    return str(p)


def lin8(v):
    return (np.asarray(v, np.float32) / 255.0) ** 2.2


# ---------------------------------------------------------------- image loading
@pytest.mark.parametrize("ctype", [2, 6, 0, 4])
def test_png_color_types_and_filters(tmp_path, ctype):
    rng = np.random.default_rng(ctype)
    nc = {0: 1, 2: 3, 4: 2, 6: 4}[ctype]
    px = rng.integers(0, 256, (7, 11, nc)).astype(np.uint8)
    img = pt.load_image(write(tmp_path, "a.png", png_bytes(px, ctype)))
    assert img.shape == (7, 11, 4)
    gray = nc <= 2
    for c in range(3):
        assert np.allclose(img[..., c], lin8(px[..., 0 if gray else c]), rtol=1e-6, atol=1e-7)
    alpha = px[..., 3] if nc == 4 else (px[..., 1] if nc == 2 else np.full(px.shape[:2], 255))
    assert np.allclose(img[..., 3], alpha / 255.0, atol=1e-7)


def test_png_palette_16bit_and_low_depth(tmp_path):
    pal = np.array([[255, 0, 0], [0, 128, 255], [10, 20, 30], [200, 200, 0]], np.uint8)
    idx = np.array([[0, 1, 2, 3, 3], [2, 1, 0, 1, 2]], np.uint8)
    img = pt.load_image(write(tmp_path, "p.png", png_bytes(idx[..., None], 3, 8, pal)))
    assert np.allclose(img[..., :3], lin8(pal[idx]), atol=1e-7)
    img = pt.load_image(write(tmp_path, "p2.png", png_bytes(idx, 3, 2, pal)))
    assert np.allclose(img[..., :3], lin8(pal[idx]), atol=1e-7)
    g4 = np.array([[0, 5, 15, 9]], np.uint8)
    img = pt.load_image(write(tmp_path, "g4.png", png_bytes(g4, 0, 4)))
    assert np.allclose(img[..., 0], lin8(g4 * 17), atol=1e-7)
    px16 = np.random.default_rng(9).integers(0, 65536, (3, 4, 3)).astype(np.uint16)
    img = pt.load_image(write(tmp_path, "s.png", png_bytes(px16, 2, 16)))
    assert np.allclose(img[..., :3], lin8(px16 >> 8), atol=1e-7)  # high byte, as an 8-bit loader keeps it


@pytest.mark.parametrize("h,w", [(1, 1), (3, 2), (9, 13), (16, 17)])
@pytest.mark.parametrize("kind", ["rgb8", "rgba16", "gray1", "gray4", "pal2", "ga8"])
def test_png_adam7_interlaced(tmp_path, kind, h, w):
    """Adam7-interlaced files (including sizes where some of the seven passes
    are empty) decode to the same image as the plain file of the same pixels,
    and to the expected values."""
    rng = np.random.default_rng(h * 31 + w)
    pal = rng.integers(0, 256, (4, 3)).astype(np.uint8)
    ctype, depth, px = {
        "rgb8": (2, 8, lambda: rng.integers(0, 256, (h, w, 3)).astype(np.uint8)),
        "rgba16": (6, 16, lambda: rng.integers(0, 65536, (h, w, 4)).astype(np.uint16)),
        "gray1": (0, 1, lambda: rng.integers(0, 2, (h, w)).astype(np.uint8)),
        "gray4": (0, 4, lambda: rng.integers(0, 16, (h, w)).astype(np.uint8)),
        "pal2": (3, 2, lambda: rng.integers(0, 4, (h, w)).astype(np.uint8)),
        "ga8": (4, 8, lambda: rng.integers(0, 256, (h, w, 2)).astype(np.uint8)),
    }[kind]
    px = px()
    palette = pal if ctype == 3 else None
    plain = pt.load_image(write(tmp_path, "a.png", png_bytes(px, ctype, depth, palette)))
    inter = pt.load_image(write(tmp_path, "i.png", png_bytes(px, ctype, depth, palette, interlace=True)))
    assert inter.shape == (h, w, 4) and np.array_equal(inter, plain)
    if kind == "rgb8":
        from PIL import Image  # an independent reader of the interlaced test file
        assert np.array_equal(np.asarray(Image.open(str(tmp_path / "i.png")).convert("RGB")), px)
        assert np.allclose(inter[..., :3], lin8(px), atol=1e-7)
    elif kind == "rgba16":
        assert np.allclose(inter[..., :3], lin8(px[..., :3] >> 8), atol=1e-7)
        assert np.allclose(inter[..., 3], (px[..., 3] >> 8) / 255.0, atol=1e-7)
    elif kind == "gray1":
        assert np.allclose(inter[..., 0], lin8(px * 255), atol=1e-7)
    elif kind == "gray4":
        assert np.allclose(inter[..., 1], lin8(px * 17), atol=1e-7)
    elif kind == "pal2":
        assert np.allclose(inter[..., :3], lin8(pal[px]), atol=1e-7)
    else:
        assert np.allclose(inter[..., 2], lin8(px[..., 0]), atol=1e-7)
        assert np.allclose(inter[..., 3], px[..., 1] / 255.0, atol=1e-7)


@pytest.mark.parametrize("rle", [False, True])
def test_radiance_hdr(tmp_path, rle):
    rng = np.random.default_rng(3)
    w = 40 if rle else 5
    rgbe = rng.integers(0, 256, (6, w, 4)).astype(np.uint8)
    rgbe[:, :8, :] = rgbe[:, :1, :]  # runs for the RLE coder
    rgbe[0, 3, 3] = 0  # a zero exponent reads black
    img = pt.load_image(write(tmp_path, "a.hdr", rgbe_bytes(rgbe, rle)))
    scale = np.where(rgbe[..., 3] > 0, np.ldexp(1.0, rgbe[..., 3].astype(np.int32) - 136), 0.0)
    assert np.allclose(img[..., :3], rgbe[..., :3] * scale[..., None], rtol=1e-7, atol=0)
    assert np.all(img[..., 3] == 1.0)


@pytest.mark.parametrize("compression,half", [(0, False), (0, True), (1, True), (2, False), (3, True), (3, False)])
def test_exr_compressions_and_pixel_types(tmp_path, compression, half):
    rng = np.random.default_rng(compression + 10 * half)
    img = rng.uniform(0, 8, (37, 13, 4)).astype(np.float32)
    img[5:20, 2:9] = 1.5  # compressible region
    if half:
        img = img.astype(np.float16).astype(np.float32)
    got = pt.load_image(write(tmp_path, "a.exr", exr_bytes(img, compression, half)))
    assert np.array_equal(got, img)


def test_exr_round_trip_with_writer(tmp_path):
    rng = np.random.default_rng(4)
    buf = rng.uniform(0, 5, (9, 6, 4)).astype(np.float32)
    p = str(tmp_path / "o.exr")
    pt.write_exr(p, buf)  # buffer row 0 = bottom of the image
    got = pt.load_image(p)
    assert np.array_equal(got[..., :3], buf[::-1, :, :3]) and np.all(got[..., 3] == 1.0)


def to8(img):
    """Linear loader output back to the 8-bit levels it came from."""
    return np.rint(np.clip(img, 0, 1) ** (1 / 2.2) * 255).astype(np.int32)


def pil_image(seed, mode, w=37, h=23):
    from PIL import Image
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    base = np.stack([xx * 255 // max(1, w - 1), yy * 255 // max(1, h - 1), (xx + yy) * 3 % 256], -1)
    px = np.clip(base + rng.integers(-20, 21, base.shape), 0, 255).astype(np.uint8)
    if mode == "RGBA":
        px = np.concatenate([px, rng.integers(0, 256, (h, w, 1)).astype(np.uint8)], -1)
    img = Image.fromarray(px if mode != "L" else px[..., 0], mode if mode in ("RGB", "RGBA", "L") else "RGB")
    return img.convert(mode) if mode == "P" else img


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "P"])
def test_bmp_written_by_pillow(tmp_path, mode):
    img = pil_image(1, mode)
    path = str(tmp_path / "a.bmp")
    img.save(path)
    from PIL import Image
    ref = np.asarray(Image.open(path).convert("RGBA"))
    got = pt.load_image(path)
    assert got.shape == ref.shape
    assert np.array_equal(to8(got[..., :3]), ref[..., :3])
    assert np.allclose(got[..., 3], ref[..., 3] / 255.0, atol=1e-7)


def bmp_bytes(px, bpp, top_down=False, masks=None):
    """Minimal BMP writer: 16-bit 565 or 32-bit with bit-field masks."""
    h, w = px.shape[:2]
    stride = (w * bpp + 31) // 32 * 4
    rows = []
    for y in (range(h) if top_down else range(h - 1, -1, -1)):
        r = bytearray()
        for x in range(w):
            v = 0
            for k, m in enumerate(masks):
                if m:
                    shift = (m & -m).bit_length() - 1
                    bits = m >> shift
                    v |= (int(px[y, x, k]) * bits // 255) << shift
            r += v.to_bytes(bpp // 8, "little")
        rows.append(bytes(r) + b"\0" * (stride - len(r)))
    data = b"".join(rows)
    hdr = struct.pack("<IiiHHIIiiII", 56, w, -h if top_down else h, 1, bpp, 3, len(data), 2835, 2835, 0, 0)
    hdr += struct.pack("<IIII", *masks)
    off = 14 + len(hdr)
    return b"BM" + struct.pack("<IHHI", off + len(data), 0, 0, off) + hdr + data


@pytest.mark.parametrize("bpp,top_down", [(16, False), (32, True)])
def test_bmp_bitfields(tmp_path, bpp, top_down):
    rng = np.random.default_rng(bpp)
    px = rng.integers(0, 256, (5, 7, 4)).astype(np.uint8)
    masks = (0xF800, 0x07E0, 0x001F, 0) if bpp == 16 else (0x00FF0000, 0x0000FF00, 0x000000FF, 0xFF000000)
    got = pt.load_image(write(tmp_path, "b.bmp", bmp_bytes(px, bpp, top_down, masks)))
    for k, m in enumerate(masks[:3]):
        bits = m >> ((m & -m).bit_length() - 1)
        q = px[..., k].astype(np.int64) * bits // 255
        assert np.array_equal(to8(got[..., k]), (q * 255 + bits // 2) // bits)
    if bpp == 32:
        assert np.allclose(got[..., 3], px[..., 3] / 255.0, atol=1e-7)


@pytest.mark.parametrize("mode,rle", [("RGB", False), ("RGBA", True), ("L", True), ("P", False), ("RGB", True)])
def test_tga_written_by_pillow(tmp_path, mode, rle):
    from PIL import Image
    img = pil_image(2, mode)
    arr = np.array(img)
    if rle:  # long runs for the run-length coder
        arr[3:9] = arr[3:4]
        img = Image.fromarray(arr, img.mode)
        if mode == "P":
            img.putpalette(pil_image(2, mode).getpalette())
    path = str(tmp_path / "a.tga")
    img.save(path, compression="tga_rle" if rle else None)
    ref = np.asarray(Image.open(path).convert("RGBA"))
    got = pt.load_image(path)
    assert np.array_equal(to8(got[..., :3]), ref[..., :3])
    assert np.allclose(got[..., 3], ref[..., 3] / 255.0, atol=1e-7)


def test_pnm_8_and_16_bit(tmp_path):
    rng = np.random.default_rng(5)
    px = rng.integers(0, 256, (6, 9, 3)).astype(np.uint8)
    got = pt.load_image(write(tmp_path, "a.ppm", b"P6\n# comment\n9 6\n255\n" + px.tobytes()))
    assert np.array_equal(to8(got[..., :3]), px)
    g16 = rng.integers(0, 65536, (4, 5)).astype(">u2")
    got = pt.load_image(write(tmp_path, "a.pgm", b"P5 5 4 65535\n" + g16.tobytes()))
    expect = (g16.astype(np.int64) * 255 + 32767) // 65535
    assert np.array_equal(to8(got[..., 0]), expect) and np.array_equal(got[..., 0], got[..., 2])


@pytest.mark.parametrize("mode,subsampling", [("L", 0), ("RGB", 0), ("RGB", 1), ("RGB", 2)])
def test_jpeg_against_pillow(tmp_path, mode, subsampling):
    """Baseline JPEG written by Pillow (libjpeg): our decode vs libjpeg's.  The
    standard leaves IDCT and upsampling rounding open, so levels may differ
    by a few; parity with stb_image's rounding is unpinned."""
    from PIL import Image
    img = pil_image(3, mode, 61, 45)
    path = str(tmp_path / "a.jpg")
    kw = {} if mode == "L" else {"subsampling": subsampling}
    img.save(path, quality=92, **kw)
    ref = np.asarray(Image.open(path).convert("RGB")).astype(np.int32)
    got = to8(pt.load_image(path)[..., :3])
    d = np.abs(got - ref)
    print(f"{mode} subsampling {subsampling}: max diff {d.max()} mean {d.mean():.3f}")
    assert d.max() <= (2 if subsampling == 0 else 4) and d.mean() <= (0.1 if subsampling == 0 else 0.4)


def test_jpeg_restart_markers_and_lossless_rejected(tmp_path):
    from PIL import Image
    img = pil_image(4, "RGB", 80, 40)
    path = str(tmp_path / "r.jpg")
    img.save(path, quality=90, subsampling=2, restart_marker_blocks=3)
    assert b"\xff\xdd" in open(path, "rb").read()
    ref = np.asarray(Image.open(path).convert("RGB")).astype(np.int32)
    got = to8(pt.load_image(path)[..., :3])
    assert np.abs(got - ref).max() <= 6
    # SOF3 (lossless) is not read: the baseline file with its frame marker changed
    data = bytearray(open(path, "rb").read())
    at = data.find(b"\xff\xc0")
    data[at + 1] = 0xC3
    with pytest.raises(pt.PupilError) as e:
        pt.load_image(write(tmp_path, "l.jpg", bytes(data)))
    assert e.value.code == -8


@pytest.mark.parametrize("mode,subsampling,restart", [("L", 0, 0), ("RGB", 0, 0), ("RGB", 2, 0), ("RGB", 1, 0),
                                                      ("RGB", 2, 2), ("L", 0, 5)])
def test_progressive_jpeg(tmp_path, mode, subsampling, restart):
    """Progressive JPEG (libjpeg's default script: DC first + refinement scans,
    spectral bands, successive approximation) re-encodes the same quantized
    coefficients as the baseline file Pillow writes with the same settings,
    so our decode of the progressive file must equal our decode of the
    baseline file exactly; against libjpeg's own decode the usual IDCT /
    upsampling rounding tolerance applies."""
    from PIL import Image
    img = pil_image(7, mode, 77, 53)
    kw = {} if mode == "L" else {"subsampling": subsampling}
    if restart:
        kw["restart_marker_blocks"] = restart
    base, prog = str(tmp_path / "b.jpg"), str(tmp_path / "p.jpg")
    img.save(base, quality=88, **kw)
    img.save(prog, quality=88, progressive=True, **kw)
    data = open(prog, "rb").read()
    assert b"\xff\xc2" in data and data.count(b"\xff\xda") > 3  # SOF2, several scans
    got_p = pt.load_image(prog)
    got_b = pt.load_image(base)
    assert np.array_equal(got_p, got_b)
    ref = np.asarray(Image.open(prog).convert("RGB")).astype(np.int32)
    d = np.abs(to8(got_p[..., :3]) - ref)
    assert d.max() <= (2 if subsampling == 0 else 4) and d.mean() <= (0.1 if subsampling == 0 else 0.4)


@pytest.mark.parametrize("interlace,transparent,size", [(False, None, (37, 23)), (True, None, (37, 23)),
                                                        (True, 5, (64, 48)), (False, 0, (300, 200))])
def test_gif_first_image(tmp_path, interlace, transparent, size):
    """GIF written by Pillow: colors equal Pillow's decode exactly (alpha 0 on
    the transparent index, as stb_image's canvas); the 300x200 low-entropy
    image fills the 4096-entry LZW table, so codes of every width and the
    full-table case are decoded."""
    from PIL import Image
    rng = np.random.default_rng(size[0] + (transparent or 0))
    w, h = size
    if w > 100:  # smooth bands: long repeats, many table entries
        idx = ((np.arange(w)[None, :] // 3 + np.arange(h)[:, None] // 5) % 64).astype(np.uint8)
        idx[rng.random((h, w)) < 0.02] = rng.integers(0, 64)
    else:
        idx = rng.integers(0, 64, (h, w)).astype(np.uint8)
    pal = rng.integers(0, 256, (64, 3)).astype(np.uint8)
    img = Image.fromarray(idx, "P")
    img.putpalette(pal.reshape(-1).tolist())
    path = str(tmp_path / "a.gif")
    kw = {"interlace": interlace}
    if transparent is not None:
        kw["transparency"] = transparent
    img.save(path, **kw)
    ref = np.asarray(Image.open(path).convert("RGBA"))
    got = pt.load_image(path)
    assert got.shape == (h, w, 4)
    alpha = np.round(got[..., 3] * 255).astype(np.int32)
    assert np.array_equal(alpha, ref[..., 3].astype(np.int32))
    opaque = ref[..., 3] == 255
    assert np.array_equal(to8(got[..., :3])[opaque], ref[..., :3][opaque])
    assert np.all(got[~opaque] == 0)
    if transparent is not None:
        assert (~opaque).any() and np.array_equal(~opaque, idx == transparent)


def packbits(row):
    """PackBits: runs of 3..128 equal bytes as (1 - n, byte), the rest as literal chunks."""
    out, i, lit = bytearray(), 0, bytearray()

    def flush():
        while lit:
            chunk = lit[:128]
            out.extend([len(chunk) - 1]); out.extend(chunk)
            del lit[:128]

    while i < len(row):
        j = i
        while j < len(row) and j - i < 128 and row[j] == row[i]:
            j += 1
        if j - i >= 3:
            flush()
            out.extend([(1 - (j - i)) & 0xFF, row[i]])
            i = j
        else:
            lit.append(row[i])
            i += 1
    flush()
    return bytes(out)


def psd_bytes(planes, depth, rle):
    """planes: (channels, h, w) samples of an RGB(A) document."""
    nch, h, w = planes.shape
    dt = ">u2" if depth == 16 else "u1"
    head = b"8BPS" + struct.pack(">H6xHIIHH", 1, nch, h, w, depth, 3)
    body = struct.pack(">I", 0) * 3  # color mode data, image resources, layers and masks: empty
    if not rle:
        return head + body + struct.pack(">H", 0) + planes.astype(dt).tobytes()
    rows = [packbits(planes[c, y].astype(dt).tobytes()) for c in range(nch) for y in range(h)]
    return head + body + struct.pack(">H", 1) + b"".join(struct.pack(">H", len(r)) for r in rows) + b"".join(rows)


@pytest.mark.parametrize("depth,rle,nch", [(8, False, 3), (8, True, 4), (16, False, 4), (16, True, 3)])
def test_psd_composite(tmp_path, depth, rle, nch):
    rng = np.random.default_rng(depth + nch + rle)
    h, w = 9, 300  # wide rows: literal chunks longer than 128 bytes
    planes = rng.integers(0, 1 << depth, (nch, h, w)).astype(np.uint16 if depth == 16 else np.uint8)
    planes[:, 2, 10:200] = planes[:, 2, 10:11]  # runs for PackBits (longer than 128)
    path = write(tmp_path, "a.psd", psd_bytes(planes, depth, rle))
    got = pt.load_image(path)
    hi = (planes >> 8) if depth == 16 else planes
    assert got.shape == (h, w, 4)
    for c in range(3):
        assert np.array_equal(to8(got[..., c]), hi[c].astype(np.int32))
    alpha = hi[3] if nch == 4 else np.full((h, w), 255)
    assert np.allclose(got[..., 3], alpha / 255.0, atol=1e-7)
    if depth == 8:
        from PIL import Image  # an independent reader of the test file
        ref = np.asarray(Image.open(path).convert("RGBA" if nch == 4 else "RGB"))
        assert np.array_equal(ref[..., :3].transpose(2, 0, 1), planes[:3])


def test_image_errors(tmp_path):
    with pytest.raises(pt.PupilError) as e:
        pt.load_image(write(tmp_path, "x.pic", b"\x53\x80\xf6\x34"))  # not Radiance
    assert e.value.code == -6
    with pytest.raises(pt.PupilError) as e:  # CMYK document
        pt.load_image(write(tmp_path, "c.psd", b"8BPS" + struct.pack(">H6xHIIHH", 1, 4, 2, 2, 8, 4)))
    assert e.value.code == -8
    with pytest.raises(pt.PupilError) as e:  # header claims 16000 x 16000 RGB, no data follows
        pt.load_image(write(tmp_path, "t.psd", b"8BPS" + struct.pack(">H6xHIIHH", 1, 3, 16000, 16000, 8, 3)
                            + struct.pack(">I", 0) * 3 + struct.pack(">H", 0)))
    assert e.value.code == -6
    gif = bytearray(b"GIF89a" + struct.pack("<HHBBB", 4, 4, 0x80, 0, 0) + bytes(6))  # 4x4 screen, 2 colors
    gif += b"\x2c" + struct.pack("<HHHHB", 0, 0, 60000, 60000, 0) + b"\x02\x02\x4c\x01\x00\x3b"
    with pytest.raises(pt.PupilError) as e:  # image descriptor far larger than the screen
        pt.load_image(write(tmp_path, "big.gif", bytes(gif)))
    assert e.value.code == -6
    with pytest.raises(pt.PupilError) as e:
        pt.load_image(write(tmp_path, "x.gif", b"GIF89a"))
    assert e.value.code == -6
    with pytest.raises(pt.PupilError) as e:
        pt.load_image(write(tmp_path, "x.jpg", b"\xff\xd8\xff"))
    assert e.value.code == -6
    with pytest.raises(pt.PupilError) as e:
        pt.load_image(write(tmp_path, "t.png", png_bytes(np.zeros((4, 4, 3), np.uint8), 2)[:40]))
    assert e.value.code == -6
    with pytest.raises(pt.PupilError) as e:
        pt.load_image(str(tmp_path / "missing.exr"))
    assert e.value.code == -5


def test_image_size_limits_before_allocation(tmp_path):
    """Every reader checks the header's size against one limit (2^24 pixels per
    side, 2^28 in all) before allocating: a 29-byte GIF with a 65535 x 65535
    screen and a 360 KB PackBits PSD claiming 30000 x 30000 x 4 x 16 bit are
    rejected at once, and PackBits row counts too short for the row width (or
    summing past the end of the file) are rejected before the planes exist."""
    gif = b"GIF89a" + struct.pack("<HHBBB", 65535, 65535, 0, 0, 0) + b"\x2c" + struct.pack("<HHHHB", 0, 0, 1, 1, 0)
    gif += b"\x02\x02\x4c\x01\x00\x3b"
    assert len(gif) == 29
    with pytest.raises(pt.PupilError) as e:
        pt.load_image(write(tmp_path, "screen.gif", gif))
    assert e.value.code == -8 and "2^28" in str(e.value)

    def psd_rle(w, h, nch, counts, depth=16, data=b""):
        head = b"8BPS" + struct.pack(">H6xHIIHH", 1, nch, h, w, depth, 3) + struct.pack(">I", 0) * 3
        return head + struct.pack(">H", 1) + struct.pack(f">{nch * h}H", *counts) + data

    big = psd_rle(30000, 30000, 4, [1] * (4 * 30000), data=bytes(4 * 30000))
    assert 350_000 < len(big) < 370_000
    with pytest.raises(pt.PupilError) as e:
        pt.load_image(write(tmp_path, "big.psd", big))
    assert e.value.code == -8
    # within the size limit: 1-byte rows cannot hold 2 x 4000 bytes each
    with pytest.raises(pt.PupilError) as e:
        pt.load_image(write(tmp_path, "short.psd", psd_rle(4000, 4000, 4, [1] * 16000, data=bytes(16000))))
    assert e.value.code == -6 and "PackBits" in str(e.value)
    # plausible counts whose sum runs past the end of the file
    with pytest.raises(pt.PupilError) as e:
        pt.load_image(write(tmp_path, "trunc.psd", psd_rle(4000, 4000, 3, [126] * 12000, data=bytes(1000))))
    assert e.value.code == -6 and "truncated" in str(e.value)


OOM_SCRIPT = """
import resource, sys
sys.path.insert(0, {root!r})
import pupiloptixlab_amd as pt
pt.lib()  # map the library before the address space is limited
vm = [int(l.split()[1]) * 1024 for l in open("/proc/self/status") if l.startswith("VmSize")][0]
resource.setrlimit(resource.RLIMIT_AS, (vm + (512 << 20), resource.RLIM_INFINITY))
try:
    pt.load_image({path!r})
except pt.PupilError as e:
    print("code", e.code)
"""


def test_host_allocation_failure_is_reported(tmp_path):
    """A reader whose pixel buffer cannot be allocated (address space limited
    to 512 MB above the interpreter's) returns PT_ERR_OOM through the C
    interface instead of terminating the process."""
    import subprocess
    import sys
    import zlib
    w = h = 16384  # 2^28 pixels: within the size limit, 4 GiB as float RGBA
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)
    raw = zlib.compress(b"")
    chunk = lambda t, d: struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", raw) + chunk(b"IEND", b"")
    path = write(tmp_path, "huge.png", png)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", OOM_SCRIPT.format(root=root, path=path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip() == "code -7", (r.stdout, r.stderr[-2000:])


# ---------------------------------------------------------------- bitmap sampling
IMG = np.arange(4 * 3 * 4, dtype=np.float32).reshape(3, 4, 4) * np.float32(0.25)


def texel(x, y):
    return IMG[y, x, :3]


def test_point_sampling_picks_floor_texel():
    for (u, v, x, y) in [(0.0, 0.0, 0, 0), (0.26, 0.34, 1, 1), (0.99, 0.99, 3, 2), (0.5, 0.5, 2, 1)]:
        assert np.array_equal(oracle.bitmap_sample(IMG, u, v, filter=0), texel(x, y))


def test_bilinear_texel_centers_and_midpoints():
    for x in range(4):
        for y in range(3):
            u, v = (x + 0.5) / 4, (y + 0.5) / 3
            assert np.allclose(oracle.bitmap_sample(IMG, u, v, wrap=1), texel(x, y), atol=1e-6)
    mid = oracle.bitmap_sample(IMG, 2.0 / 4, 0.5 / 3, wrap=1)  # halfway between texels (1,0) and (2,0)
    assert np.allclose(mid, 0.5 * (texel(1, 0) + texel(2, 0)), atol=1e-6)
    q = oracle.bitmap_sample(IMG, 1.75 / 4, 1.0 / 3, wrap=1)  # quarter across, halfway down
    want = 0.75 * (0.5 * texel(1, 0) + 0.5 * texel(1, 1)) + 0.25 * (0.5 * texel(2, 0) + 0.5 * texel(2, 1))
    assert np.allclose(q, want, atol=1e-5)


def test_address_modes():
    # repeat: integer shifts change nothing; the left edge blends with the right column
    for u, v in [(0.3, 0.6), (0.9, 0.1)]:
        assert np.allclose(oracle.bitmap_sample(IMG, u + 2, v - 1), oracle.bitmap_sample(IMG, u, v), atol=1e-5)
    edge = oracle.bitmap_sample(IMG, 0.0, 0.5 / 3)
    assert np.allclose(edge, 0.5 * (texel(0, 0) + texel(3, 0)), atol=1e-6)
    # clamp: outside reads the edge texel
    assert np.allclose(oracle.bitmap_sample(IMG, -0.7, 0.5 / 3, filter=0, wrap=1), texel(0, 0))
    assert np.allclose(oracle.bitmap_sample(IMG, 1.7, 2.5 / 3, filter=0, wrap=1), texel(3, 2))
    # mirror: [1, 2) reflects [0, 1)
    assert np.allclose(oracle.bitmap_sample(IMG, 1.1, 0.2, filter=0, wrap=2), oracle.bitmap_sample(IMG, 0.9, 0.2, filter=0, wrap=2))
    # border: red (1, 0, 0) outside
    assert np.allclose(oracle.bitmap_sample(IMG, -0.5, 0.5, filter=0, wrap=3), [1, 0, 0])
    assert np.allclose(oracle.bitmap_sample(IMG, 0.6, 0.5, filter=0, wrap=3), texel(2, 1))
    # non-finite coordinates read coordinate 0
    assert np.allclose(oracle.bitmap_sample(IMG, np.nan, 0.5 / 3, filter=0), texel(0, 0))


# ---------------------------------------------------------------- XML
TEX_XML = """<scene version="3.0.0">
  <integrator type="path"><integer name="max_depth" value="{depth}"/></integrator>
  <sensor type="perspective"><float name="fov" value="{fov}"/>
    <transform name="to_world"><lookat origin="{cam}" target="0, 0, 0" up="0, 1, 0"/></transform>
    <film type="hdrfilm"><integer name="width" value="{w}"/><integer name="height" value="{h}"/></film>
  </sensor>
  <shape type="rectangle">
    <transform name="to_world"><scale x="50" y="50"/><rotate x="1" angle="-90"/></transform>
    <bsdf type="diffuse">{floor}</bsdf>
  </shape>
  {extra}
  <emitter type="envmap"><string name="filename" value="{env}"/><float name="scale" value="{scale}"/>{env_tr}</emitter>
</scene>"""


def env_scene(tmp_path, env_img, w=2, h=2, fov=0.1, depth=2, scale=1.0, floor='<rgb name="reflectance" value="0.5"/>',
              extra="", cam="0, 3, 0.3", env_tr=""):
    envp = write(tmp_path, "env.exr", exr_bytes(env_img, 3, False))
    text = TEX_XML.format(w=w, h=h, fov=fov, depth=depth, scale=scale, floor=floor, extra=extra, cam=cam, env=envp,
                          env_tr=env_tr)
    return pt.HostScene.load_xml(write(tmp_path, "scene.xml", text))


def test_xml_bitmap_texture_and_envmap(tmp_path):
    px = np.random.default_rng(5).integers(0, 256, (4, 8, 3)).astype(np.uint8)
    write(tmp_path, "tex.png", png_bytes(px, 2))
    floor = """<texture type="bitmap" name="reflectance"><string name="filename" value="tex.png"/>
        <string name="filter_type" value="bilinear"/><string name="wrap_mode" value="mirror"/>
        <transform name="to_uv"><scale x="2" y="3"/><translate x="0.5" y="0.25"/></transform></texture>"""
    extra = """<shape type="sphere"><bsdf type="diffuse"><texture type="bitmap" name="reflectance">
        <string name="filename" value="tex.png"/></texture></bsdf></shape>"""
    env = np.ones((8, 16, 4), np.float32)
    s = env_scene(tmp_path, env, floor=floor, extra=extra, scale=2.5,
                  env_tr='<transform name="to_world"><rotate y="1" angle="90"/></transform>')
    d = s.desc
    bms = s.bitmaps()
    assert d.num_bitmaps == 3 and d.env_type == pt.ENV_MAP and d.env_scale == 2.5
    (w0, h0, f0, m0, a0), (w1, h1, f1, m1, a1), (w2, h2, f2, m2, a2) = bms
    assert (w0, h0, f0, m0) == (8, 4, pt.FILTER_LINEAR, pt.WRAP_MIRROR)
    assert (w1, h1, f1, m1) == (8, 4, pt.FILTER_POINT, pt.WRAP_REPEAT)  # no filter_type: point
    assert (w2, h2, f2, m2) == (16, 8, pt.FILTER_LINEAR, pt.WRAP_REPEAT)
    assert np.allclose(a0[..., :3], lin8(px), atol=1e-7) and a0.ctypes.data == a1.ctypes.data  # one image
    floor_mat = [m for m in s.materials() if m.tex[0].type == pt.TEX_BITMAP][0]
    assert list(floor_mat.tex[0].uv_row0) == pytest.approx([2, 0, 0, 0.5])
    assert list(floor_mat.tex[0].uv_row1) == pytest.approx([0, 3, 0, 0.25])
    assert np.allclose(np.array(d.env_to_world).reshape(3, 3), [[0, 0, 1], [0, 1, 0], [-1, 0, 0]], atol=1e-6)


# ---------------------------------------------------------------- env-map integrator
def filtered_env(img, phi, theta):
    """Bilinear, repeat-addressed radiance of an (h, w, 4) map at angles."""
    h, w = img.shape[:2]
    u, v = phi / (2 * np.pi), theta / np.pi
    x, y = u * w - 0.5, v * h - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    a, c = x - x0, y - y0
    x0, y0 = x0.astype(int), y0.astype(int)
    lum = img[..., :3] @ np.array([0.2126, 0.7152, 0.0722])
    L = lambda i, j: lum[np.mod(j, h), np.mod(i, w)]
    return ((1 - a) * (1 - c) * L(x0, y0) + a * (1 - c) * L(x0 + 1, y0) + (1 - a) * c * L(x0, y0 + 1)
            + a * c * L(x0 + 1, y0 + 1))


def floor_irradiance(img, scale=1.0, n=2048):
    """E = int L(w) max(0, cos) dw for a floor with normal +y in the map frame."""
    th = (np.arange(n) + 0.5) * np.pi / n
    ph = (np.arange(2 * n) + 0.5) * np.pi / n
    P, T = np.meshgrid(ph, th)
    Lw = filtered_env(img, P, T)
    cos = np.cos(T)  # local +y is cos(theta)
    return scale * float(np.sum(Lw * np.maximum(cos, 0) * np.sin(T)) * (np.pi / n) ** 2)


def luminance_image(env, rgb):
    return float(np.asarray(rgb) @ np.array([0.2126, 0.7152, 0.0722]))


@pytest.mark.parametrize("kind", ["uniform", "random", "sun"])
def test_envmap_floor_closed_form_exact_sampler(tmp_path, kind):
    rng = np.random.default_rng(11)
    h, w = 16, 32
    if kind == "uniform":
        env = np.full((h, w, 4), 0.7, np.float32)
    elif kind == "random":
        env = rng.uniform(0, 3, (h, w, 4)).astype(np.float32)
    else:
        env = np.full((h, w, 4), 0.02, np.float32)
        env[3, 7, :3] = [400.0, 300.0, 200.0]  # a bright texel above the horizon
    env[..., 3] = 1.0
    gray = np.repeat((env[..., :3] @ np.array([0.2126, 0.7152, 0.0722], np.float32))[..., None], 3, axis=2)
    env_g = np.concatenate([gray, env[..., 3:]], axis=2).astype(np.float32)  # gray map: radiance = luminance
    s = env_scene(tmp_path, env_g, scale=1.5)
    want = 0.5 / np.pi * floor_irradiance(env_g, 1.5)
    out = oracle.render(s.desc, spp=16384, max_depth=2, mis_mode=pt.MIS_CONSISTENT)
    got = float(out["accum"][..., :3].mean())
    assert got == pytest.approx(want, rel=0.02), (kind, got, want)


def test_envmap_reference_sampler_is_close(tmp_path):
    rng = np.random.default_rng(12)
    env = rng.uniform(0.5, 1.5, (32, 64, 4)).astype(np.float32)
    s = env_scene(tmp_path, env)
    want_lum = 0.5 / np.pi * floor_irradiance(env)
    out = oracle.render(s.desc, spp=8192, max_depth=2, mis_mode=pt.MIS_REFERENCE)
    got = float(out["accum"][..., :3].mean(axis=(0, 1)) @ np.array([0.2126, 0.7152, 0.0722]))
    assert got == pytest.approx(want_lum, rel=0.1)


def test_envmap_rotation_and_primary_view(tmp_path):
    """Rotating the map by 180 degrees about x puts the sky below the floor:
    the floor then sees the map's lower half."""
    env = np.zeros((8, 16, 4), np.float32)
    env[:4, :, :3] = 2.0  # upper hemisphere bright
    env[4:, :, :3] = 0.25
    env[..., 3] = 1
    up = env_scene(tmp_path, env)
    flip = env_scene(tmp_path, env, env_tr='<transform name="to_world"><rotate x="1" angle="180"/></transform>')
    a = float(oracle.render(up.desc, spp=4096, max_depth=2, mis_mode=1)["accum"][..., :3].mean())
    b = float(oracle.render(flip.desc, spp=4096, max_depth=2, mis_mode=1)["accum"][..., :3].mean())
    assert a == pytest.approx(0.5 / np.pi * floor_irradiance(env), rel=0.03)
    assert b == pytest.approx(0.5 / np.pi * floor_irradiance(env[::-1].copy()), rel=0.03)
    # a camera looking straight up sees the map radiance (at the pole the
    # repeat-wrapped bilinear filter blends the first and last rows)
    env[7, :, :3] = 2.0
    sky = env_scene(tmp_path, env, cam="0, 1, 0", depth=1)
    d = sky.desc
    import ctypes as C
    c2w = np.array([[1, 0, 0, 0], [0, 0, -1, 1], [0, 1, 0, 0], [0, 0, 0, 1]], np.float32)  # looks along +y
    for i in range(16):
        d.camera.camera_to_world[i] = float(c2w.reshape(-1)[i])
    out = oracle.render(d, spp=1, max_depth=1)
    assert np.allclose(out["accum"][..., :3], 2.0, rtol=1e-5)


# ---------------------------------------------------------------- GPU parity
REL_TOL, PRIM_TOL = 1e-3, 1e-4


def textured_scene(tmp_path):
    rng = np.random.default_rng(21)
    write(tmp_path, "a.png", png_bytes(rng.integers(0, 256, (16, 24, 3)).astype(np.uint8), 2))
    write(tmp_path, "b.hdr", rgbe_bytes(rng.integers(100, 140, (8, 8, 4)).astype(np.uint8), False))
    floor = """<texture type="bitmap" name="reflectance"><string name="filename" value="a.png"/>
        <string name="filter_type" value="bilinear"/><transform name="to_uv"><scale x="7" y="7"/></transform></texture>"""
    extra = "".join(f"""<shape type="sphere"><transform name="to_world"><scale value="0.4"/>
        <translate value="{-1 + i}, 0.4, 0"/></transform>
        <bsdf type="{'roughplastic' if i == 1 else 'diffuse'}"><texture type="bitmap" name="{'diffuse_reflectance' if i == 1 else 'reflectance'}">
        <string name="filename" value="{'b.hdr' if i == 2 else 'a.png'}"/><string name="wrap_mode" value="{['clamp', 'mirror', 'repeat'][i]}"/>
        {'<string name="filter_type" value="bilinear"/>' if i != 0 else ''}</texture></bsdf></shape>""" for i in range(3))
    extra += """<shape type="rectangle"><transform name="to_world"><scale value="0.3"/><rotate x="1" angle="90"/>
        <translate value="0, 2, 0"/></transform><bsdf type="diffuse"/>
        <emitter type="area"><texture type="bitmap" name="radiance"><string name="filename" value="b.hdr"/></texture></emitter></shape>"""
    env = rng.uniform(0.0, 2.0, (16, 32, 4)).astype(np.float32)
    env[2, 5, :3] = 80.0
    return env_scene(tmp_path, env, w=96, h=64, fov=50, depth=6, floor=floor, extra=extra, cam="0, 1.6, 3.2",
                     env_tr='<transform name="to_world"><rotate y="1" angle="30"/></transform>')


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [pt.MIS_REFERENCE, pt.MIS_CONSISTENT])
def test_gpu_textures_and_envmap_parity(tracer, tmp_path, mode):
    print(f"case: bitmap textures + env map 96x64 8spp depth 6 mis={mode}", flush=True)
    s = textured_scene(tmp_path)
    tracer.set_scene(s)
    tracer.build_bvh()
    tracer.render(spp=8, mis_mode=mode)
    acc, prim = tracer.read(pt.BUF_ACCUM), tracer.read(pt.BUF_PRIM_ID)
    ref = oracle.render(s.desc, spp=8, mis_mode=mode)
    g, r = acc[..., :3].astype(np.float64), ref["accum"][..., :3].astype(np.float64)
    err = np.abs(g - r) / np.maximum(np.abs(r), 1e-6 / REL_TOL)
    outliers = float(np.mean(err.max(axis=-1) > REL_TOL))
    exact = float(np.mean(np.all(g == r, axis=-1)))
    print(f"prim mismatch {np.mean(prim != ref['prim']):.2e} outliers {outliers:.2e} bit-identical {exact:.4f}")
    assert float(np.mean(prim != ref["prim"])) <= PRIM_TOL and outliers == 0.0
    assert np.isfinite(acc).all()


@pytest.mark.gpu
def test_gpu_envmap_closed_form(tracer, tmp_path):
    print("case: random env map over a floor, exact sampler, 2x2 16384 spp", flush=True)
    env = np.random.default_rng(13).uniform(0, 3, (16, 32, 4)).astype(np.float32)
    env[..., 3] = 1
    s = env_scene(tmp_path, env)
    tracer.set_scene(s)
    tracer.build_bvh()
    tracer.render(spp=16384, max_depth=2, mis_mode=pt.MIS_CONSISTENT)
    got = tracer.read(pt.BUF_ACCUM)[..., :3].mean(axis=(0, 1)) @ np.array([0.2126, 0.7152, 0.0722])
    assert got == pytest.approx(0.5 / np.pi * floor_irradiance(env), rel=0.02)


@pytest.mark.gpu
def test_gpu_batch_textured_views(tracer, tmp_path):
    print("case: batch of a textured env-map view and a Cornell view vs pt_render", flush=True)
    scenes = [(textured_scene(tmp_path), 3), (pt.HostScene.generate(pt.SCENE_CORNELL, width=64, height=64), 2)]
    ref = []
    for sc, spp in scenes:
        tracer.set_scene(sc)
        tracer.build_bvh()
        tracer.render(spp=spp, mis_mode=pt.MIS_REFERENCE, aov=False)
        ref.append(tracer.read(pt.BUF_ACCUM))
    tracer.batch_clear()
    for sc, spp in scenes:
        tracer.batch_add_view(sc, spp)
    tracer.batch_render(mis_mode=pt.MIS_REFERENCE)
    for v, r in enumerate(ref):
        assert np.array_equal(tracer.batch_read(v), r), f"view {v}"
    tracer.batch_clear()
