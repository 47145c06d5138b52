"""The `drainage` program and its stage classes on the GPU.

- the fused device loop and the stage-by-stage plugin path (--staged) give
  the same counter raster, and both equal the oracle's Monte Carlo loop on the
  same synthetic inputs, seed and realization indices;
- a run stopped at a resume point and continued with --resume equals the
  uninterrupted run;
- frames output writes one decodable JPEG per iteration;
- ToByteArray on the device follows the reference formula
  rintf(255 * ((v - min) / max)) (ToByteArrayAlgorithmGPUCUDA.cu:22-84).
Every case prints its name before it touches the GPU.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cageo_amd", "bin", "drainage")
W, H = 320, 256


def announce(*what):
    print("CASE", *what, flush=True)


def run_cli(cwd, *args):
    if not os.path.exists(CLI):
        pytest.fail("drainage CLI not built")
    r = subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


@pytest.fixture(scope="module")
def inputs(tmp_path_factory, torch_cuda):
    d = tmp_path_factory.mktemp("dem")
    announce("cli synth", W, H)
    run_cli(str(d), "--synth", f"{W},{H},5", "--dem-path", ".", "-i", "0")
    return d


def read_bin(path, shape=(H, W)):
    return np.fromfile(path, np.float32).reshape(shape)


def oracle_mc(d, n, seed, std, model=1, pr=60.0, first=0):
    dem = read_bin(os.path.join(d, "dem10.bin"))
    st = read_bin(os.path.join(d, "streams.bin"))
    import cageo_amd as ca
    r = oracle.filter_radius(pr, 10.0)
    f = ca.create_filter(model, pr, 10.0, r)  # the taps the program uses (host, no device)
    mc = np.zeros((H, W), np.float32)
    for k in range(first, first + n):
        oracle.lib().or_realization(dem, st, W, H, f, r, model, 0.0, std, seed, k, -9999.0,
                                    -9999.0, 2, mc)
    return mc


def test_cli_fused_staged_and_oracle(inputs):
    d = str(inputs)
    announce("cli fused")
    run_cli(d, "--dem-path", ".", "-i", "4", "--seed", "99", "--error-standard-deviation", "1.8",
            "-o", "fused.bin")
    announce("cli staged")
    run_cli(d, "--dem-path", ".", "-i", "4", "--seed", "99", "--error-standard-deviation", "1.8",
            "-o", "sub/staged.bin", "--staged")  # stem only: written to the working directory
    fused, staged = read_bin(os.path.join(d, "fused.bin")), read_bin(os.path.join(d, "staged.bin"))
    assert np.array_equal(fused, staged)
    assert fused.max() <= 4 and fused.sum() > 0
    assert np.array_equal(fused, oracle_mc(d, 4, 99, 1.8))


def test_cli_exponential(inputs):
    d = str(inputs)
    args = ["--dem-path", ".", "-i", "1", "--seed", "7", "--filter-type", "Exponential",
            "--practical-range", "200", "--error-standard-deviation", "2.5"]
    announce("cli exponential fused/staged")
    run_cli(d, *args, "-o", "e1.bin")
    run_cli(d, *args, "-o", "e2.bin", "--staged")
    a, b = read_bin(os.path.join(d, "e1.bin")), read_bin(os.path.join(d, "e2.bin"))
    assert np.array_equal(a, b)
    assert np.array_equal(a, oracle_mc(d, 1, 7, 2.5, model=0, pr=200.0))


def test_cli_resume(inputs):
    d = str(inputs)
    common = ["--dem-path", ".", "--seed", "31", "--error-standard-deviation", "1.8"]
    announce("cli resume")
    run_cli(d, *common, "-i", "5", "-o", "whole.bin")
    run_cli(d, *common, "-i", "2", "--checkpoint", "ck")
    r = run_cli(d, *common, "-i", "5", "--resume", "ck", "-o", "resumed.bin", "--checkpoint-every",
                "2", "--checkpoint", "ck2")
    assert "Resuming at realization 2" in r.stdout
    assert np.array_equal(read_bin(os.path.join(d, "whole.bin")),
                          read_bin(os.path.join(d, "resumed.bin")))
    # a resume point made with other settings is refused
    bad = subprocess.run([CLI, "--dem-path", ".", "--seed", "31", "-i", "5", "--resume", "ck",
                          "--practical-range", "90"], cwd=d, capture_output=True, text=True,
                         timeout=120)
    assert bad.returncode == 1 and "practical_range" in bad.stdout


def test_cli_frames_jpeg_and_timing(inputs):
    Image = pytest.importorskip("PIL.Image")
    d = str(inputs)
    announce("cli frames")
    r = run_cli(d, "--dem-path", ".", "-i", "2", "--seed", "3", "--output-option", "frames",
                "-o", "fr.jpg", "--timing")
    for k in range(2):
        img = np.asarray(Image.open(os.path.join(d, f"fr_{k:04d}.jpg")))
        assert img.shape == (H, W)
    assert "Timer summary" in r.stdout and "Pit filling" in r.stdout
    assert "[TIMING] Flow tracing" in r.stderr


def reference_bytes(x):
    x = x.astype(np.float32)
    mn, mx = x.min(), x.max()
    with np.errstate(invalid="ignore", divide="ignore"):
        v = np.rint(np.float32(255.0) * ((x - mn) / mx))
    q = np.where(np.isfinite(v), v, 0).astype(np.int64)
    return (q & 0xFF).astype(np.uint8)


@pytest.mark.parametrize("case", ["counts", "with_nodata", "zeros"])
def test_to_byte_array(torch_cuda, case):
    import cageo_amd as ca
    torch = torch_cuda
    r = np.random.default_rng(3)
    x = r.integers(0, 50, (H, W)).astype(np.float32)
    if case == "with_nodata":
        x[r.random((H, W)) < 0.05] = -1.0  # counter nodata: (v - min) / max exceeds 1
    if case == "zeros":
        x[:] = 0.0  # 0 / 0
    announce("to_byte_array", case)
    with ca.Context(W, H) as ctx:
        out = torch.zeros((H, W), dtype=torch.uint8, device="cuda")
        ctx.to_byte_array(out, torch.from_numpy(x).cuda())
        got = out.cpu().numpy()
    assert np.array_equal(got, reference_bytes(x))


def test_cli_surrogate_accept_and_fallback(inputs, torch_cuda):
    """--surrogate: an accepted prediction is written as probability x
    iterations (the C-ABI prediction of the same inputs); a rejected one
    leaves the output to the Monte Carlo, equal to a run without the
    surrogate.  Acceptance and rejection are forced by the L1 target (10:
    any map passes; 0: only an exact one would).  A windowed run refines
    its failing windows with a Monte Carlo instead of rejecting."""
    import cageo_amd as ca
    from cageo_amd import surrogate as sg
    torch = torch_cuda
    d = str(inputs)
    announce("cli surrogate accept / fallback")
    common = ["--dem-path", ".", "-i", "6", "--seed", "9", "--error-standard-deviation", "1.5",
              "--surrogate", sg.DEFAULT_WEIGHTS]
    r = run_cli(d, *common, "--surrogate-l1", "10", "-o", "sur_ok.bin")
    assert "accepted" in r.stdout, r.stdout
    got = read_bin(os.path.join(d, "sur_ok.bin"))
    dem = torch.from_numpy(read_bin(os.path.join(d, "dem10.bin"))).cuda()
    st = torch.from_numpy(read_bin(os.path.join(d, "streams.bin"))).cuda()
    p = torch.empty((H, W), dtype=torch.float32, device="cuda")
    with ca.Surrogate(sg.DEFAULT_WEIGHTS) as s:
        s.predict(dem, st, -9999.0, -9999.0, 1.5, 60.0, p)
    torch.cuda.synchronize()
    assert np.array_equal(got, p.cpu().numpy() * np.float32(6))
    r = run_cli(d, *common, "--surrogate-l1", "0", "-o", "sur_no.bin")
    assert "rejected" in r.stdout, r.stdout
    run_cli(d, "--dem-path", ".", "-i", "6", "--seed", "9", "--error-standard-deviation", "1.5",
            "-o", "plain.bin")
    assert np.array_equal(read_bin(os.path.join(d, "sur_no.bin")),
                          read_bin(os.path.join(d, "plain.bin")))
    # windows of 64 cells, all failing at target 0: every window takes the
    # refinement's frequency (count / 16) x iterations
    r = run_cli(d, *common, "--surrogate-l1", "0", "--surrogate-window", "64",
                "--surrogate-refine", "16", "-o", "sur_win.bin")
    assert "failed" in r.stdout and "refined" in r.stdout, r.stdout
    win = read_bin(os.path.join(d, "sur_win.bin"))
    counts = win / np.float32(6) * np.float32(16)
    assert np.allclose(counts, np.round(counts), atol=1e-3) and counts.max() <= 16


def test_cli_fast_convolution(inputs, torch_cuda):
    """--fast-convolution with the exponential filter runs and gives counts
    in [0, iterations] whose total is within 1 % of the exact order's."""
    d = str(inputs)
    announce("cli fast convolution")
    base = ["--dem-path", ".", "-i", "3", "--seed", "4", "--filter-type", "Exponential",
            "--practical-range", "200", "--error-standard-deviation", "2.5"]
    run_cli(d, *base, "-o", "exact.bin")
    run_cli(d, *base, "--fast-convolution", "-o", "fast.bin")
    a, b = read_bin(os.path.join(d, "exact.bin")), read_bin(os.path.join(d, "fast.bin"))
    assert b.min() >= 0 and b.max() <= 3
    assert abs(float(b.sum()) - float(a.sum())) <= 0.01 * float(a.sum())


def test_cli_batch(tmp_path, torch_cuda):
    """--batch: three tiles of different sizes and error models listed in a
    file run together (graph mode, tuned); tile t's counter raster equals
    the fused single-tile run of that DEM with seed + t, and the oracle's
    loop.  A tile line may omit the error model (then the options give it)."""
    d = str(tmp_path)
    announce("cli batch")
    sizes = [(200, 130, 5), (96, 160, 6), (128, 128, 7)]
    for k, (w, h, s) in enumerate(sizes):
        run_cli(d, "--synth", f"{w},{h},{s}", "--dem-path", ".", "--dem", f"t{k}",
                "--stream", f"t{k}s", "-i", "0")
    with open(os.path.join(d, "tiles.txt"), "w") as f:
        f.write("# dem stream [error std, practical range, filter]\n")
        f.write("t0 t0s 1.5 60 Gaussian\n")
        f.write("t1 t1s 2.5 150 Exponential\n")
        f.write("t2 t2s   # error model from the options\n")
    r = run_cli(d, "--batch", "tiles.txt", "--batch-tune", "--dem-path", ".", "-i", "3",
                "--seed", "40", "--error-standard-deviation", "1.0", "--practical-range", "80",
                "-o", "bt.bin")
    assert "3 tiles x 3" in r.stdout, r.stdout
    models = [(1.5, 60.0, 1), (2.5, 150.0, 0), (1.0, 80.0, 1)]
    for k, ((w, h, _), (std, pr, model)) in enumerate(zip(sizes, models)):
        got = np.fromfile(os.path.join(d, f"bt_t{k}.bin"), np.float32).reshape(h, w)
        ftype = "Gaussian" if model == 1 else "Exponential"
        run_cli(d, "--dem-path", ".", "--dem", f"t{k}", "--stream", f"t{k}s", "-i", "3",
                "--seed", str(40 + k), "--error-standard-deviation", str(std),
                "--practical-range", str(pr), "--filter-type", ftype, "-o", f"one{k}.bin")
        one = np.fromfile(os.path.join(d, f"one{k}.bin"), np.float32).reshape(h, w)
        assert np.array_equal(got, one), f"tile {k}"
        dem = np.fromfile(os.path.join(d, f"t{k}.bin"), np.float32).reshape(h, w)
        st = np.fromfile(os.path.join(d, f"t{k}s.bin"), np.float32).reshape(h, w)
        import cageo_amd as ca
        rad = oracle.filter_radius(pr, 10.0)
        taps = ca.create_filter(model, pr, 10.0, rad)
        ref = np.zeros((h, w), np.float32)
        for i in range(3):
            oracle.lib().or_realization(dem, st, w, h, taps, rad, model, 0.0, std, 40 + k, i,
                                        -9999.0, -9999.0, 2, ref)
        assert np.array_equal(got, ref), f"tile {k} vs oracle"
