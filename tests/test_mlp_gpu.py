"""GPU MLP kernels (BF16 MFMA) against the CPU FP64 MLP oracle.

The oracle rounds to bf16 where the GPU feeds the matrix cores, so the
remaining differences are FP32-vs-FP64 accumulation and rare one-ulp bf16
flips of an activation: forward within 2e-3 of the output scale, gradients
within 1e-2 relative Frobenius norm per tensor.  The fused inference kernel
meets the same forward bound (its K order inside an MFMA differs from the
training GEMMs', so FP32 sums and hence rare bf16 roundings differ), and one
Adam step from zero moments moves every trained weight by lr * sign(grad)."""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def setup(seed=0, n=200):
    import torch
    from sysmat import mlp
    rng = np.random.default_rng(seed)
    p = omlp.init(seed)
    for l in range(1, 5):
        p[omlp.OFF[f"b{l}"]:omlp.OFF[f"b{l}"] + 256] = rng.normal(0, 0.2, 256)
    p[omlp.OFF["b5"]] = 0.3
    p32 = p.astype(np.float32)
    x = rng.normal(0, 1, (n, 12)).astype(np.float32)
    net = mlp.Mlp(params=p32)
    return net, p32.astype(np.float64), x, torch.from_numpy(x).cuda()


def test_forward_matches_oracle(lib):
    import torch
    for n in (200, 64, 1000):
        net, p, x, xd = setup(1, n)
        y = net.forward(xd).cpu().numpy()
        yo, _ = omlp.forward(p, x.astype(np.float64), emulate_bf16=True)
        yo_exact, _ = omlp.forward(p, x.astype(np.float64), emulate_bf16=False)
        scale = np.abs(yo).max()
        err = np.abs(y - yo).max() / scale
        err_exact = np.abs(y - yo_exact).max() / scale
        print(f"n={n}: forward err vs bf16 oracle {err:.2e}, vs fp64 oracle {err_exact:.2e}", flush=True)
        assert err < 2e-3
        assert err_exact < 5e-2


def test_backward_matches_oracle(lib):
    import torch
    net, p, x, xd = setup(2, 300)
    rng = np.random.default_rng(7)
    dy = rng.normal(0, 1, 300).astype(np.float32)
    net.forward(xd)
    g = net.backward(torch.from_numpy(dy).cuda()).cpu().numpy().astype(np.float64)
    _, cache = omlp.forward(p, x.astype(np.float64), emulate_bf16=True)
    go = omlp.backward(p, cache, dy.astype(np.float64), emulate_bf16=True)
    gu, gou = omlp.unpack(g), omlp.unpack(go)
    for k in gou:
        rel = np.linalg.norm(gu[k] - gou[k]) / max(np.linalg.norm(gou[k]), 1e-30)
        print(f"grad {k}: rel err {rel:.2e}", flush=True)
        assert rel < 1e-2, k
    assert np.all(gu["W1"][:, 12:] == 0)


def test_fused_inference_equals_training_forward(lib):
    from sysmat import mlp
    net, p, x, xd = setup(3, 777)
    net2 = mlp.Mlp(seed=5)
    packed = mlp.pack_members([net, net2])
    y = mlp.infer(packed, 2, xd).cpu().numpy()
    y0 = net.forward(xd).cpu().numpy()
    y1 = net2.forward(xd).cpu().numpy()
    p2 = net2.params.cpu().numpy().astype(np.float64)
    for a, b, pp in ((y[0], y0, p), (y[1], y1, p2)):
        yo, _ = omlp.forward(pp, x.astype(np.float64), emulate_bf16=True)
        err = np.abs(a - b).max() / np.abs(b).max()
        err_o = np.abs(a - yo).max() / np.abs(yo).max()
        print(f"fused vs training forward: {err:.2e}, vs bf16 oracle {err_o:.2e}", flush=True)
        assert err < 2e-3 and err_o < 2e-3


def test_adam_first_step_is_sign_of_gradient(lib):
    import torch
    net, p, x, xd = setup(4, 256)
    t = torch.from_numpy(np.random.default_rng(1).normal(0, 1, 256).astype(np.float32)).cuda()
    y = net.forward(xd)
    dy = 2.0 * (y - t) / 256.0
    g = net.backward(dy).cpu().numpy()
    p0 = net.params.cpu().numpy()
    lr = 1e-3
    net.train_step(xd, t, lr=lr)
    d = net.params.cpu().numpy() - p0
    big = np.abs(g) > 1e-4
    assert big.sum() > 1000
    gb = g[big].astype(np.float64)
    want = -lr * np.sqrt(1 - 0.999) / (1 - 0.9) * (0.1 * gb) / (np.sqrt(0.001 * gb * gb) + 1e-8)
    np.testing.assert_allclose(d[big], want, rtol=2e-4, atol=2e-8)
    assert np.all(np.abs(np.abs(want) - lr) <= 3.2e-3 * lr)
    assert np.all(d[np.abs(g) == 0] == 0)


def test_training_reduces_loss(lib):
    import torch
    from sysmat import mlp
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (8192, 12)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2] + 0.5 * x[:, 3] ** 2).astype(np.float32)
    xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    net = mlp.Mlp(seed=11)
    first = None
    for it in range(400):
        idx = torch.randint(0, 8192, (1024,), device="cuda")
        loss = net.train_step(xd[idx].contiguous(), td[idx].contiguous(), lr=2e-3, want_loss=(it % 50 == 0))
        if loss is not None:
            v = float(loss.item())
            print(f"step {it}: loss {v:.4e}", flush=True)
            first = v if first is None else first
    final = float(((net.forward(xd) - td) ** 2).mean().item())
    print(f"final mse {final:.3e} (start {first:.3e}, var {t.var():.3e})", flush=True)
    assert final < 0.02 * t.var()


M64 = (1 << 64) - 1


def _mix64(z):
    """splitmix64 finaliser on Python ints (the device sampler's hash)."""
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _batch_indices(seed, step, n, B):
    base = _mix64(seed ^ ((0x632BE59BD9B4E019 * step) & M64))
    return np.array([_mix64((base + b) & M64) % n for b in range(B)], dtype=np.int64)


def _pairs(n, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 12)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2]).astype(np.float32)
    w = rng.uniform(0.2, 2.0, n).astype(np.float32)
    return x, t, w


def _normalised_weights(wsel):
    """The device's weight normalisation in its own float32 order: per block of 256 rows the
    wave sums by xor butterflies, then the four waves in order; the block sums in order;
    then w * (1 / total)."""
    B = len(wsel)
    nb = (B + 255) // 256
    wpad = np.zeros(nb * 256, np.float32)
    wpad[:B] = wsel
    tot = np.float32(0.0)
    for blk in wpad.reshape(nb, 4, 64):
        waves = blk.copy()
        for o in (32, 16, 8, 4, 2, 1):
            waves = (waves + waves[:, np.arange(64) ^ o]).astype(np.float32)
        bs = np.float32(0.0)
        for q in range(4):
            bs = np.float32(bs + waves[q, 0])
        tot = np.float32(tot + bs)
    return (wsel * (np.float32(1.0) / tot)).astype(np.float32)


def _cosine_lr(lr, floor, it, steps):
    """max(lr 0.5 (1 + cosf(pi it / steps)), floor) in float32 with the C library's cosf,
    as the multi-step entry point evaluates it on the host."""
    import ctypes
    import ctypes.util
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.cosf.restype = ctypes.c_float
    libm.cosf.argtypes = [ctypes.c_float]
    f = np.float32
    arg = f(f(f(3.14159265358979) * f(it)) / f(steps))
    v = f(f(f(lr) * f(0.5)) * f(f(1.0) + f(libm.cosf(arg))))
    return float(max(v, f(floor)))


def test_train_run_equals_stepwise_training_on_the_same_batches(lib):
    """sm_mlp_train_run draws batch row b of step s at mix64(mix64(seed ^ K s) + b) mod n on
    the device (restated here in Python), normalises the batch weights and follows a cosine
    schedule; the same batches, weights and learning rates fed to the one-step entry point
    give bit-identical parameters and loss."""
    import torch
    from sysmat import mlp
    n, B, steps, lr, floor, seed = 5000, 256, 24, 2e-3, 2e-5, 77
    x, t, w = _pairs(n)
    xd, td, wd = (torch.from_numpy(a).cuda() for a in (x, t, w))
    a = mlp.Mlp(seed=5)
    b = mlp.Mlp(seed=5)
    la = a.train_run(xd, td, wd, steps=steps, lr=lr, lr_floor=floor, seed=seed, batch=B, want_loss=True)
    for it in range(steps):
        idx = _batch_indices(seed, it + 1, n, B)
        wb = _normalised_weights(w[idx])
        ii = torch.from_numpy(idx).cuda()
        lb = b.train_step(xd[ii].contiguous(), td[ii].contiguous(), torch.from_numpy(wb).cuda(),
                          lr=_cosine_lr(lr, floor, it, steps), batch=B, want_loss=(it == steps - 1))
    pa, pb = a.params.cpu().numpy(), b.params.cpu().numpy()
    err = np.abs(pa - pb).max() / np.abs(pb).max()
    print(f"train_run vs stepwise: max param diff {err:.2e} of the largest parameter; "
          f"loss {float(la.item()):.8e} vs {float(lb.item()):.8e}", flush=True)
    assert a.step == b.step == steps
    assert np.array_equal(pa, pb)
    assert float(la.item()) == float(lb.item())


def test_train_run_is_deterministic_and_learns(lib):
    import torch
    from sysmat import mlp
    n = 8192
    x, t, w = _pairs(n, seed=4)
    xd, td, wd = (torch.from_numpy(a).cuda() for a in (x, t, w))
    nets = [mlp.Mlp(seed=11) for _ in range(2)]
    for net in nets:
        net.train_run(xd, td, wd, steps=400, lr=2e-3, lr_floor=2e-5, seed=9, batch=1024)
    assert torch.equal(nets[0].params, nets[1].params)
    final = float(((nets[0].forward(xd) - td) ** 2).mean().item())
    print(f"train_run final mse {final:.3e} (var {t.var():.3e})", flush=True)
    assert final < 0.02 * t.var()


def test_members_trained_together_equal_members_trained_alone(lib):
    import torch
    from sysmat import mlp
    n = 6000
    x, t, w = _pairs(n, seed=6)
    xd, td, wd = (torch.from_numpy(a).cuda() for a in (x, t, w))
    together = [mlp.Mlp(seed=21 + m) for m in range(2)]
    mlp.train_members(together, xd, td, wd, steps=40, lr=2e-3, lr_floor=2e-5, seeds=[3, 4], batch=512)
    for m in range(2):
        alone = mlp.Mlp(seed=21 + m)
        alone.train_run(xd, td, wd, steps=40, lr=2e-3, lr_floor=2e-5, seed=3 + m, batch=512)
        assert torch.equal(alone.params, together[m].params)
