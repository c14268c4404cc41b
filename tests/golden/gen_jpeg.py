"""Writes the JPEG fixtures of tests/test_jpeg.py: synthetic images encoded
by libjpeg-turbo through Pillow (the JPEG codec OpenCV also builds on; the
reference reads its images with cv::imread(IMREAD_GRAYSCALE),
HPM-MVS/HPM.cpp:492), and for each the grey image libjpeg decodes
(grayscale output: the luminance component, integer inverse DCT), as PGM,
and the colour image it decodes (RGB output, default upsampling, as
cv::imread(IMREAD_COLOR) asks for in RunFusion, main.cpp:523), as PPM.
Progressive files (spectral selection and successive approximation, as
libjpeg's default progression script writes them) have expected images too.

    python tests/golden/gen_jpeg.py
"""
import io
import os

import numpy as np
from PIL import Image, features

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jpeg")


def content(h, w, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    base = 128 + 60 * np.sin(x / 3.7) * np.cos(y / 5.3) + 0.8 * (x - y)
    base[(x > w / 3) & (x < w / 2) & (y > h / 4)] = 20  # a sharp-edged block
    base += rng.normal(0, 6, (h, w))
    return np.clip(base, 0, 255).astype(np.uint8)


def rgb(h, w, seed):
    return np.stack([content(h, w, seed), content(h, w, seed + 1)[:, ::-1], content(h, w, seed + 2)[::-1]], axis=-1)


CASES = [
    ("grey_q75_37x23", lambda: Image.fromarray(content(23, 37, 1)), dict(quality=75)),
    ("ycc420_q90_101x77", lambda: Image.fromarray(rgb(77, 101, 2)), dict(quality=90, subsampling=2)),
    ("ycc444_q50_64x48", lambda: Image.fromarray(rgb(48, 64, 3)), dict(quality=50, subsampling=0)),
    ("ycc422_q95_restart_70x45", lambda: Image.fromarray(rgb(45, 70, 4)),
     dict(quality=95, subsampling=1, restart_marker_blocks=5)),
    ("ycc420_optimized_120x90", lambda: Image.fromarray(rgb(90, 120, 5)), dict(quality=85, subsampling=2, optimize=True)),
    ("grey_restart_rows_q100_33x17", lambda: Image.fromarray(content(17, 33, 6)),
     dict(quality=100, restart_marker_rows=1)),
    ("progressive_48x32", lambda: Image.fromarray(rgb(32, 48, 7)), dict(quality=80, progressive=True)),
    ("progressive_grey_q90_53x41", lambda: Image.fromarray(content(41, 53, 8)), dict(quality=90, progressive=True)),
    ("progressive_ycc420_optimized_101x77", lambda: Image.fromarray(rgb(77, 101, 9)),
     dict(quality=70, subsampling=2, progressive=True, optimize=True)),
    ("progressive_ycc444_restart_q95_66x50", lambda: Image.fromarray(rgb(50, 66, 10)),
     dict(quality=95, subsampling=0, progressive=True, restart_marker_blocks=3)),
]


def main():
    os.makedirs(HERE, exist_ok=True)
    print("libjpeg", features.version("jpg"), "turbo", features.check_feature("libjpeg_turbo"))
    for name, make, opts in CASES:
        buf = io.BytesIO()
        make().save(buf, "JPEG", **opts)
        data = buf.getvalue()
        with open(os.path.join(HERE, name + ".jpg"), "wb") as f:
            f.write(data)
        im = Image.open(io.BytesIO(data))
        im.draft("L", im.size)
        g = np.asarray(im.convert("L") if im.mode != "L" else im)
        with open(os.path.join(HERE, name + ".pgm"), "wb") as f:
            f.write(b"P5\n%d %d\n255\n" % (g.shape[1], g.shape[0]))
            f.write(g.astype(np.uint8).tobytes())
        c = np.asarray(Image.open(io.BytesIO(data)).convert("RGB"))
        with open(os.path.join(HERE, name + ".ppm"), "wb") as f:
            f.write(b"P6\n%d %d\n255\n" % (c.shape[1], c.shape[0]))
            f.write(c.astype(np.uint8).tobytes())
        print(name, len(data), "bytes", g.shape)


if __name__ == "__main__":
    main()
