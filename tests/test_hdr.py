"""Radiance .hdr reader (alpine_amd/hdr.py) for environment maps: decoding pinned by the
RGBE format's closed form (mantissa x 2^(e - 136), e = 0 black) on files written here in
both scanline encodings, and error paths.  The reference reads these files with stb_image
(loaders/hdr_loader.cpp:10-26), which is absent, so equality with stb itself is unpinned
beyond the format's definition."""
import numpy as np
import pytest

from alpine_amd import hdr, simple

HEADER = b"#?RADIANCE\n# made by the tests\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n"


def _rle_channel(vals):
    """Encode one channel of a scanline with runs of >= 3 equal bytes and literals."""
    out, i, n = bytearray(), 0, len(vals)
    while i < n:
        j = i
        while j < n and j - i < 127 and vals[j] == vals[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + (j - i), vals[i]])
            i = j
            continue
        j = i
        while j < n and j - i < 128 and not (j + 2 < n and vals[j] == vals[j + 1] == vals[j + 2]):
            j += 1
        out += bytes([j - i]) + bytes(vals[i:j])
        i = j
    return bytes(out)


def write_hdr(path, rgbe, rle):
    h, w, _ = rgbe.shape
    data = bytearray(HEADER + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        if rle:
            data += bytes([2, 2, w >> 8, w & 255])
            for c in range(4):
                data += _rle_channel(list(rgbe[y, :, c]))
        else:
            data += rgbe[y].tobytes()
    open(path, "wb").write(bytes(data))


def _random_rgbe(h, w, seed):
    rng = np.random.default_rng(seed)
    rgbe = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    rgbe[..., 3] = rng.integers(100, 160, (h, w))
    rgbe[: h // 3, : w // 2] = rgbe[0, 0]           # long runs
    rgbe[h // 2, :, 3] = 0                          # a black scanline
    return rgbe


@pytest.mark.parametrize("rle", [False, True])
@pytest.mark.parametrize("shape", [(7, 5), (19, 300)])
def test_decodes_both_encodings(tmp_path, rle, shape):
    h, w = shape
    rgbe = _random_rgbe(h, w, seed=h * w)
    p = str(tmp_path / "env.hdr")
    write_hdr(p, rgbe, rle and w >= 8)
    img = hdr.read_hdr(p)
    assert img.shape == (h, w, 3) and img.dtype == np.float32
    e = rgbe[..., 3].astype(np.float64)
    want = np.where(e[..., None] > 0, rgbe[..., :3] * np.exp2(e[..., None] - 136), 0.0)
    np.testing.assert_array_equal(img, want.astype(np.float32))
    assert (img[h // 2] == 0).all()


def test_closed_forms_and_extremes():
    px = np.array([[[128, 64, 255, 129], [1, 2, 3, 0], [255, 255, 255, 1], [255, 0, 1, 255]]],
                  np.uint8)
    f = hdr.rgbe_to_float(px)[0]
    np.testing.assert_array_equal(f[0], [1.0, 0.5, 255 / 128])
    np.testing.assert_array_equal(f[1], [0, 0, 0])
    assert f[2, 0] == np.float32(255 * 2.0 ** -135)       # subnormal in float32
    assert f[3, 0] == np.float32(255 * 2.0 ** 119) and f[3, 1] == 0


def test_flat_file_with_rle_capable_width(tmp_path):
    """A picture of width >= 8 whose first scanline has no run-length marker is flat."""
    rgbe = _random_rgbe(4, 16, 1)
    rgbe[0, 0, :2] = (200, 7)
    p = str(tmp_path / "flat.hdr")
    write_hdr(p, rgbe, rle=False)
    np.testing.assert_array_equal(hdr.read_hdr(p), hdr.rgbe_to_float(rgbe))


@pytest.mark.parametrize("bad", ["magic", "format", "orient", "trunc", "width", "runlen"])
def test_rejects_malformed(tmp_path, bad):
    rgbe = _random_rgbe(3, 12, 2)
    p = str(tmp_path / "bad.hdr")
    write_hdr(p, rgbe, rle=True)
    data = open(p, "rb").read()
    if bad == "magic":
        data = data.replace(b"#?RADIANCE", b"#?NOTHING!")
    elif bad == "format":
        data = data.replace(b"FORMAT=32-bit_rle_rgbe", b"FORMAT=32-bit_rle_xyze")
    elif bad == "orient":
        data = data.replace(b"-Y 3 +X 12", b"+Y 3 +X 12")
    elif bad == "trunc":
        data = data[:-5]
    elif bad == "width":
        data = data.replace(bytes([2, 2, 0, 12]), bytes([2, 2, 0, 13]), 1)
    else:
        i = data.index(bytes([2, 2, 0, 12])) + 4
        data = data[:i] + bytes([128 + 100]) + data[i + 1:]
    open(p, "wb").write(data)
    with pytest.raises(hdr.HdrError):
        hdr.read_hdr(p)


def test_simple_rejects_missing_environment_before_the_gpu(tmp_path, capsys):
    (tmp_path / "a.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    rc = simple.main(["-i", str(tmp_path / "a.obj"), "-o", str(tmp_path / "o.ppm"), "--spp", "1",
                      "--accelerator", "gpu", "--lightSampler", "bvh",
                      "--environment", str(tmp_path / "none.hdr")])
    assert rc == 1 and "ERROR" in capsys.readouterr().out


@pytest.mark.gpu
def test_simple_renders_with_an_environment_map(tmp_path):
    """A uniform blue environment replaces the white background on every miss."""
    print("[gpu] simple cli with environment map", flush=True)
    (tmp_path / "a.obj").write_text("v -0.2 -0.2 0\nv 0.2 -0.2 0\nv 0 0.2 0\nf 1 2 3\n")
    rgbe = np.zeros((8, 16, 4), np.uint8)
    rgbe[...] = (16, 32, 255, 128)    # mantissa x 2^-8: (0.0625, 0.125, 0.996)
    env = str(tmp_path / "env.hdr")
    write_hdr(env, rgbe, rle=True)
    out = str(tmp_path / "o.pfm")
    assert simple.main(["-i", str(tmp_path / "a.obj"), "-o", out, "--spp", "2", "--accelerator",
                        "gpu", "--lightSampler", "bvh", "--environment", env]) == 0
    img = simple.read_pfm(out) if hasattr(simple, "read_pfm") else None
    if img is None:
        raw = open(out, "rb").read()
        img = np.frombuffer(raw[-256 * 256 * 12:], np.float32).reshape(256, 256, 3)
    corner = img[0, 0]
    np.testing.assert_allclose(corner, [16 / 256, 32 / 256, 255 / 256], rtol=1e-6)
