"""Surrogate training behind the C ABI (cageo_trainer_*, cageo_surrogate_train).

The trainer's forward and backward passes run the same layer kernels as the
Python U-Net (cageo_amd/surrogate.py), whose gradients the float64 oracle
checks in test_gpu_surrogate.py; here the C trainer's loss and gradients
must equal the Python network's on the same weights and batch (up to the
order of the float atomics in the weight and bias gradients), a run split
into pieces must draw the batches and schedule of one run, training must
lower the loss, and the weights file it writes must load as the
surrogate's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def announce(*what):
    print("CASE", *what, flush=True)


@pytest.fixture(scope="module")
def env(torch_cuda):
    import cageo_amd as ca
    from cageo_amd import surrogate as sg
    ca.lib()
    return ca, sg, torch_cuda


def batch(rng, n, h, w):
    x4 = np.zeros((n, h, w, 4), np.float32)
    x4[..., 0] = rng.normal(0.0, 1.0, (n, h, w))
    x4[..., 1] = (rng.random((n, h, w)) < 0.1).astype(np.float32)
    x4[..., 2] = rng.uniform(0.2, 1.0, (n, 1, 1))
    x4[..., 3] = rng.uniform(0.15, 1.0, (n, 1, 1))
    y = rng.random((n, h, w)).astype(np.float32) * 0.3
    return x4, y


def test_trainer_gradients_equal_python_network(env, tmp_path):
    ca, sg, torch = env
    announce("trainer gradients vs Python U-Net, base 32, 2 x 32 x 32")
    path = str(tmp_path / "init.bin")
    sg.save_params(sg.init_params(32, seed=5), path)
    params = sg.load_params(path)  # bf16-rounded convolution weights, as the file holds them
    net = sg.UNet(torch, base=32, params=params)
    rng = np.random.default_rng(7)
    x4, y = batch(rng, 2, 32, 32)
    xp = np.zeros(x4.shape[:3] + (sg.IN_PAD,), np.float32)
    xp[..., :4] = x4
    net.zero_grad()
    net.forward(torch.from_numpy(xp).cuda())
    loss_py = net.backward(torch.from_numpy(y).cuda())
    g_py = net.grads_numpy()
    with ca.Trainer(32, init_weights=path, stream=torch.cuda.current_stream()) as tr:
        loss_c = tr.gradients(torch.from_numpy(x4).cuda(), torch.from_numpy(y).cuda())
        assert abs(loss_c - loss_py) <= 1e-5 * abs(loss_py), (loss_c, loss_py)
        for name, g in g_py.items():
            gc = tr.read(name, g.shape, grad=True)
            scale = float(np.abs(g).max()) or 1.0
            err = float(np.abs(gc - g).max()) / scale
            assert err <= 2e-3, (name, err)


def test_trainer_pieces_equal_one_run_and_loss_falls(env):
    ca, sg, torch = env
    rng = np.random.default_rng(11)
    n, h = 24, 32
    x4, _ = batch(rng, n, h, h)
    # a target the inputs determine: high next to stream cells, low elsewhere
    s = x4[..., 1]
    near = np.maximum.reduce([np.roll(np.roll(s, dy, 1), dx, 2)
                              for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
    y = (0.05 + 0.85 * near).astype(np.float32)
    X, Y = torch.from_numpy(x4).cuda(), torch.from_numpy(y).cuda()
    stream = torch.cuda.current_stream()
    announce("cageo_surrogate_train: 2 pieces of 40 steps vs one run of 80")
    with ca.Trainer(32, seed=3, stream=stream) as a, ca.Trainer(32, seed=3, stream=stream) as b:
        la = np.concatenate([a.train(X, Y, 4, 2e-3, 0.02, 0, 40, 80, 99),
                             a.train(X, Y, 4, 2e-3, 0.02, 40, 40, 80, 99)])
        lb = b.train(X, Y, 4, 2e-3, 0.02, 0, 80, 80, 99)
        # the same batches and schedule: the first steps agree up to the order
        # of the float atomics, which then makes the runs drift apart slowly
        assert np.allclose(la[:3], lb[:3], rtol=1e-4), (la[:3], lb[:3])
        assert abs(la[-10:].mean() - lb[-10:].mean()) <= 0.05 * lb[-10:].mean()
        assert lb[-10:].mean() < 0.8 * lb[:5].mean(), (lb[:5], lb[-10:])


def test_trainer_weights_file_loads_as_the_surrogate(env, tmp_path):
    ca, sg, torch = env
    announce("trainer save -> surrogate weights")
    path = str(tmp_path / "trained.bin")
    rng = np.random.default_rng(13)
    x4, y = batch(rng, 4, 32, 32)
    with ca.Trainer(32, seed=4, stream=torch.cuda.current_stream()) as tr:
        tr.step(torch.from_numpy(x4).cuda(), torch.from_numpy(y).cuda(), 1e-3)
        tr.save(path)
        w = tr.read("dec2a.w", (9, 64, 192))
        hb = tr.read("head.b", (1,))
    p = sg.load_params(path)
    assert np.array_equal(p["dec2a.w"], sg.bf16_round(w))
    assert np.array_equal(p["head.b"], hb)
    with ca.Surrogate(path) as s:  # the library accepts it
        dem = rng.normal(100.0, 5.0, (40, 48)).astype(np.float32)
        st = np.full(dem.shape, -9999.0, np.float32)
        out = torch.empty(dem.shape, dtype=torch.float32, device="cuda")
        s.predict(torch.from_numpy(dem).cuda(), torch.from_numpy(st).cuda(), -9999.0, -9999.0,
                  1.0, 60.0, out)
        q = out.cpu().numpy()
        assert np.isfinite(q).all() and (q >= 0).all() and (q <= 1).all()
