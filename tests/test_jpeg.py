"""JPEG input (hpm_read_image_grey): the grey image the reference's
cv::imread(path, IMREAD_GRAYSCALE) gives (HPM-MVS/HPM.cpp:492), i.e.
libjpeg's grayscale decode, on fixtures written by tests/golden/gen_jpeg.py
(libjpeg-turbo through Pillow): grey and YCbCr files, 4:2:0 / 4:2:2 / 4:4:4,
odd sizes, restart markers, optimised Huffman tables, qualities 50-100,
progressive files (spectral selection, successive approximation).
Bar: byte-exact. Other processes (lossless, arithmetic) are rejected."""
import glob
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "jpeg")
CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.pgm")))


def _pgm(path):
    data = open(path, "rb").read()
    parts = data.split(b"\n", 3)
    w, h = map(int, parts[1].split())
    return np.frombuffer(parts[3], np.uint8).reshape(h, w)


@pytest.mark.parametrize("name", CASES)
def test_decode_matches_libjpeg(hpm, name):
    got = hpm.read_image_grey(os.path.join(HERE, name + ".jpg"))
    want = _pgm(os.path.join(HERE, name + ".pgm"))
    assert got.shape == want.shape
    diff = np.abs(got.astype(int) - want.astype(int))
    assert diff.max() == 0, f"{name}: {int((diff > 0).sum())} pixels differ, max {diff.max()}"


def test_lossless_process_rejected(hpm, tmp_path):
    """A frame of an unsupported process (the lossless SOF3 marker written
    over a baseline file's SOF0) is an error, not a wrong image."""
    data = bytearray(open(os.path.join(HERE, "grey_q75_37x23.jpg"), "rb").read())
    i = data.find(b"\xff\xc0")
    assert i > 0
    data[i + 1] = 0xC3
    p = tmp_path / "lossless.jpg"
    p.write_bytes(bytes(data))
    with pytest.raises(hpm.HpmError):
        hpm.read_image_grey(str(p))
    with pytest.raises(hpm.HpmError):
        hpm.read_image_rgb(str(p))


def test_pgm_still_read(hpm, tmp_path):
    img = (np.arange(35, dtype=np.uint8) * 7).reshape(5, 7)
    p = tmp_path / "a.pgm"
    p.write_bytes(b"P5\n7 5\n255\n" + img.tobytes())
    assert np.array_equal(hpm.read_image_grey(str(p)), img.astype(np.float32))


def test_decode_random_vs_libjpeg(hpm, tmp_path):
    """40 random images, sizes, qualities, subsamplings and restart settings,
    encoded and decoded by libjpeg-turbo through Pillow when it is present."""
    Image = pytest.importorskip("PIL.Image")
    import io

    rng = np.random.default_rng(11)
    for k in range(40):
        h, w = int(rng.integers(1, 90)), int(rng.integers(1, 120))
        y, x = np.mgrid[0:h, 0:w]
        chans = [np.clip(128 + 70 * np.sin(x / rng.uniform(1, 9) + c) * np.cos(y / rng.uniform(1, 9))
                         + rng.normal(0, rng.uniform(0, 40), (h, w)), 0, 255).astype(np.uint8) for c in range(3)]
        grey = rng.random() < 0.3
        im = Image.fromarray(chans[0] if grey else np.stack(chans, -1))
        opts = dict(quality=int(rng.integers(5, 101)))
        if not grey:
            opts["subsampling"] = int(rng.integers(0, 3))
        if rng.random() < 0.3:
            opts["restart_marker_blocks"] = int(rng.integers(1, 7))
        if rng.random() < 0.3:
            opts["optimize"] = True
        buf = io.BytesIO()
        im.save(buf, "JPEG", **opts)
        path = tmp_path / f"r{k}.jpg"
        path.write_bytes(buf.getvalue())
        ref = Image.open(io.BytesIO(buf.getvalue()))
        ref.draft("L", ref.size)
        want = np.asarray(ref)
        got = hpm.read_image_grey(str(path))
        assert np.array_equal(got, want.astype(np.float32)), (k, h, w, opts)


def _ppm(path):
    data = open(path, "rb").read()
    parts = data.split(b"\n", 3)
    w, h = map(int, parts[1].split())
    return np.frombuffer(parts[3], np.uint8).reshape(h, w, 3)


COLOUR_CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.ppm")))


@pytest.mark.parametrize("name", COLOUR_CASES)
def test_colour_decode_matches_libjpeg(hpm, name):
    """cv::imread(IMREAD_COLOR) of RunFusion (main.cpp:523): libjpeg's RGB
    output (default upsampling of 2x1 / 2x2 chroma, fixed-point YCbCr
    conversion), byte-exact; grey files give the grey value per channel."""
    got = hpm.read_image_rgb(os.path.join(HERE, name + ".jpg"))
    want = _ppm(os.path.join(HERE, name + ".ppm"))
    assert got.shape == want.shape
    diff = np.abs(got.astype(int) - want.astype(int))
    assert diff.max() == 0, f"{name}: {int((diff > 0).any(-1).sum())} pixels differ, max {diff.max()}"


def test_colour_decode_random_files(hpm, tmp_path):
    """40 random colour images in every subsampling Pillow writes, odd
    sizes, qualities 30-100, sequential and progressive, against
    libjpeg-turbo's RGB decode."""
    Image = pytest.importorskip("PIL.Image")
    import io

    rng = np.random.default_rng(12)
    for k in range(40):
        h, w = int(rng.integers(1, 60)), int(rng.integers(1, 60))
        img = (rng.integers(0, 256, (h, w, 3)) * rng.uniform(0, 1, (1, 1, 3))).astype(np.uint8)
        buf = io.BytesIO()
        Image.fromarray(img).save(buf, "JPEG", quality=int(rng.integers(30, 101)), subsampling=int(k % 3),
                                  progressive=bool(k % 2), optimize=bool(k % 4 == 1))
        path = tmp_path / f"{k}.jpg"
        path.write_bytes(buf.getvalue())
        want = np.asarray(Image.open(io.BytesIO(buf.getvalue())).convert("RGB"))
        got = hpm.read_image_rgb(str(path))
        assert np.array_equal(got, want), (k, h, w, k % 3)
