"""Host-side IO and the `drainage` command line, on the CPU (no device calls).

- .hdr parsing follows utils.cpp:58-112 (tags, byte order, unknown tag = error);
- the JPEG writer's output decodes with an independent decoder (Pillow) to the
  written image, within the quantisation error of the chosen quality;
- the CLI's option handling mirrors main.cpp:329-416: --help, the output
  extension rule, unknown filter/output options, and --cpu (the reference's
  own CPU classes are not part of this build) all end before any device use.
"""
import os
import subprocess

import numpy as np
import pytest

import cageo_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cageo_amd", "bin", "drainage")


def write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return str(path)


def test_header_fields(tmp_path):
    p = write(tmp_path / "a.hdr", "ncols 4096\nnrows 2048\nxllcorner 350000.5\nyllcorner 6700000\n"
                                  "cellsize 10\nNODATA_value -9999\nbyteorder LSBFIRST\n")
    h = ca.read_header(p)
    assert h == {"ncols": 4096, "nrows": 2048, "xllcorner": 350000.5, "yllcorner": 6700000.0,
                 "cellsize": 10.0, "NODATA_value": -9999.0}
    # upper-case BYTEORDER and any tag order are accepted
    p2 = write(tmp_path / "b.hdr", "BYTEORDER MSBFIRST nrows 3 ncols 5 cellsize 2.5\n")
    h2 = ca.read_header(p2)
    assert (h2["ncols"], h2["nrows"], h2["cellsize"]) == (5, 3, 2.5)


@pytest.mark.parametrize("text", ["ncols 4\nfoo 1\n", "ncols 4\nbyteorder MIDDLE\n", "ncols x\n"])
def test_header_errors(tmp_path, text):
    p = write(tmp_path / "bad.hdr", text)
    with pytest.raises(ca.CageoError):
        ca.read_header(p)
    with pytest.raises(ca.CageoError):
        ca.read_header(str(tmp_path / "missing.hdr"))


@pytest.mark.parametrize("shape", [(100, 77), (8, 8), (1, 1), (17, 9), (64, 200)])
def test_jpeg_decodes_to_the_image(tmp_path, shape):
    Image = pytest.importorskip("PIL.Image")
    h, w = shape
    yy, xx = np.mgrid[0:h, 0:w]
    img = (127 + 100 * np.sin(xx / 9.0) * np.cos(yy / 13.0)).astype(np.uint8)
    img[h // 3: h // 2, w // 4: w // 2] = 255
    p = str(tmp_path / "x.jpg")
    ca.write_jpeg_gray(p, img, 100)
    dec = np.asarray(Image.open(p))
    assert dec.shape == shape and dec.dtype == np.uint8
    # quality 100 = unit quantisers: only the coefficient rounding is lost
    assert np.abs(dec.astype(int) - img).max() <= 2
    # at quality 75 the loss is that of libjpeg's encoder (same Annex K
    # tables, same IJG quality scaling) on the same image
    ca.write_jpeg_gray(p, img, 75)
    ours = np.abs(np.asarray(Image.open(p)).astype(int) - img).mean()
    pq = str(tmp_path / "pil.jpg")
    Image.fromarray(img).save(pq, quality=75, subsampling=0)
    theirs = np.abs(np.asarray(Image.open(pq)).astype(int) - img).mean()
    assert ours <= 1.25 * theirs + 0.25


def test_jpeg_constant_and_noise(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    p = str(tmp_path / "c.jpg")
    for v in (0, 37, 255):
        ca.write_jpeg_gray(p, np.full((24, 40), v, np.uint8), 90)
        assert (np.asarray(Image.open(p)) == v).all()
    noise = np.random.default_rng(1).integers(0, 256, (64, 64)).astype(np.uint8)
    ca.write_jpeg_gray(p, noise, 100)  # stresses long AC runs and 0xFF stuffing
    assert np.abs(np.asarray(Image.open(p)).astype(int) - noise).max() <= 2


def cli(*args, cwd=None):
    if not os.path.exists(CLI):
        pytest.skip("drainage CLI not built (run __graft_entry__.build())")
    return subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=60)


def test_cli_help():
    r = cli("--help")
    assert r.returncode == 0
    for opt in ["--dem", "--stream", "--dem-path", "--stream-path", "--filter-type",
                "--practical-range", "--error-mean", "--error-standard-deviation", "--iterations",
                "--gpu", "--cpu", "--output-file", "--output-option", "--seed"]:
        assert opt in r.stdout


@pytest.mark.parametrize("args,message", [
    (["-o", "result.png"], "Unknown output extension provided: .png"),
    (["--filter-type", "Spherical"], "Unknown filter type provided: Spherical"),
    (["--output-option", "sometimes"], "Unknown output option provided: sometimes"),
    (["--practical-range", "sixty"], "Bad value for --practical-range"),
    (["--bogus"], "unrecognised option '--bogus'"),
    (["-i"], "the required argument for option '--iterations' is missing"),
    (["--cpu"], "not built into this program"),
])
def test_cli_option_errors(args, message):
    r = cli(*args)
    assert r.returncode == 1
    assert message in r.stdout
