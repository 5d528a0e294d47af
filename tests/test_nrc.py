"""Neural radiance cache: the device forward / backward against a float64 restatement.

CPU: the restatement's own gradients are pinned by central finite differences (with the loss
denominator held fixed, as the loss defines it), on MLP weights of every layer and on table
entries of coarse and fine levels. GPU: inference and gradients of the device kernels against the
restatement with the device's FP16 rounding points; a training run lowers the loss.
"""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from oracle import nrc_ref as R

BBOX = ((0.0, 0.0, 0.0), 64.0)


def level_info(log2_table=14, base=16.0, top=2048.0):
    b = np.exp((np.log(top) - np.log(base)) / 15)
    scale = np.floor(base * b ** np.arange(16)).astype(np.float32)
    dense = ((scale.astype(np.float64) + 1) ** 3 <= 2 ** log2_table).astype(np.uint32)
    return scale, dense


def make_samples(rng, n, valid_frac=0.9):
    s = np.zeros(n, A.PATH_SAMPLE_DTYPE)
    s["pos"] = rng.uniform(1.0, 63.0, (n, 3)).astype(np.float32)
    face = rng.integers(0, 6, n)
    s["flags"] = (np.where(rng.uniform(size=n) < valid_frac, A.native.PS_X1_HIT, 0) | (face << 8)).astype(np.uint32)
    d = rng.normal(size=(n, 3))
    s["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    s["target"] = rng.uniform(0.0, 3.0, (n, 3)).astype(np.float32)
    return s


def random_params(rng, log2_table):
    grid = (rng.uniform(-0.5, 0.5, 16 * (1 << log2_table) * 2)).astype(np.float32)
    mlp = np.zeros(R.MLP_PARAMS, np.float32)
    w1, w2, w3, w4 = (rng.uniform(-1, 1, sz) for sz in ((64, 48), (64, 64), (64, 64), (16, 64)))
    w1[:, 39:] = 0.0
    w4[3:] = 0.0
    lim = lambda fan: np.sqrt(6.0 / fan)
    mlp[R.OFF_W1:R.OFF_W2] = (w1 * lim(39)).ravel()
    mlp[R.OFF_W2:R.OFF_W3] = (w2 * lim(64)).ravel()
    mlp[R.OFF_W3:R.OFF_W4] = (w3 * lim(64)).ravel()
    mlp[R.OFF_W4:] = (w4 * lim(64)).ravel()
    return grid, mlp


def test_restatement_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    log2 = 10
    scale, dense = level_info(log2)
    s = make_samples(rng, 40)
    grid, mlp = random_params(rng, log2)
    grid = grid.astype(np.float64)
    mlp = mlp.astype(np.float64)
    y, _, g, tg = R.forward_backward(s, grid, mlp, scale, dense, *BBOX, f16=False)
    den = y ** 2 + 0.01
    f = lambda gr, ml: R.loss_fixed_den(s, gr, ml, scale, dense, *BBOX, den)
    h = 1e-6
    for idx in [R.OFF_W1 + 5, R.OFF_W1 + 64 * 0 + 33, R.OFF_W2 + 700, R.OFF_W3 + 2000, R.OFF_W4 + 64 + 9,
                R.OFF_W4 + 2 * 64 + 63]:
        mp, mm = mlp.copy(), mlp.copy()
        mp[idx] += h
        mm[idx] -= h
        fd = (f(grid, mp) - f(grid, mm)) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-5 * max(1.0, abs(fd)), (idx, fd, g[idx])
    flat = tg.ravel()
    for idx in np.argsort(-np.abs(flat))[:6]:  # table entries the samples touch most
        gp, gm = grid.copy(), grid.copy()
        gp[idx] += h
        gm[idx] -= h
        fd = (f(gp, mlp) - f(gm, mlp)) / (2 * h)
        assert abs(fd - flat[idx]) <= 1e-5 * max(1.0, abs(fd)), (idx, fd, flat[idx])


def test_level_resolutions_and_dense_levels():
    scale, dense = level_info(19)
    assert scale[0] == 16 and scale[-1] == 2048 and np.all(np.diff(scale) > 0)
    assert dense[0] == 1 and dense[-1] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("log2", [14, 16])
def test_gpu_infer_and_gradients_match_restatement(ctx, log2):
    """2^14 entries per level: the table gradient is accumulated per level in LDS and reduced
    (k_nrc_scatter, k_nrc_reduce); 2^16: global atomics. Both add to the gradient buffers."""
    import torch

    rng = np.random.default_rng(1)
    nrc = A.Nrc(A.NrcParams.default(log2_table=log2))
    scale, dense = nrc.scale, nrc.dense
    assert np.array_equal(scale, level_info(log2)[0]) and np.array_equal(dense, level_info(log2)[1])
    grid, mlp = random_params(rng, log2)
    nrc.set(grid, mlp)
    n = 1000
    s = make_samples(rng, n)
    d = torch.from_numpy(s.view(np.uint8).copy()).cuda()
    out = torch.zeros(n * 4, dtype=torch.float32, device="cuda")
    nrc.infer(d.data_ptr(), n, out.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(n, 4)
    y, loss, g, tg = R.forward_backward(s, grid, mlp, scale, dense, *BBOX, f16=True)
    print(f"[nrc] |y| max {np.abs(y).max():.3f}", flush=True)
    assert np.allclose(got[:, :3], y, rtol=2e-3, atol=2e-3), np.abs(got[:, :3] - y).max()
    assert np.array_equal(got[:, 3], ((s["flags"] & 4) != 0).astype(np.float32))
    # inference visits the queries in Morton order of their positions: a permuted batch (and a
    # second, smaller call reusing the sort workspace) gives the permuted outputs bit for bit
    perm = rng.permutation(n)
    for m in (n, n // 3):
        dp = torch.from_numpy(s[perm[:m]].view(np.uint8).copy()).cuda()
        outp = torch.zeros(m * 4, dtype=torch.float32, device="cuda")
        nrc.infer(dp.data_ptr(), m, outp.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(outp.cpu().numpy().reshape(m, 4).view(np.uint32), got[perm[:m]].view(np.uint32))
    nrc.zero_grads()
    nrc.loss(reset=True)
    nrc.backward(d.data_ptr(), n)
    _, _, gg, mg = nrc.get()
    dl = nrc.loss(reset=True)
    assert abs(dl - loss) <= 1e-3 * abs(loss)
    for name, a, b in (("W1", mg[R.OFF_W1:R.OFF_W2], g[R.OFF_W1:R.OFF_W2]), ("W2", mg[R.OFF_W2:R.OFF_W3], g[R.OFF_W2:R.OFF_W3]),
                       ("W3", mg[R.OFF_W3:R.OFF_W4], g[R.OFF_W3:R.OFF_W4]), ("W4", mg[R.OFF_W4:], g[R.OFF_W4:]),
                       ("table", gg, tg.ravel())):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        print(f"[nrc] {name}: rel err {err:.2e}", flush=True)
        assert err < 1e-2, name
    # a second backward adds to the gradients
    nrc.backward(d.data_ptr(), n)
    _, _, gg2, mg2 = nrc.get()
    nrc.loss(reset=True)
    for name, a, b in (("table x2", gg2, 2.0 * gg), ("MLP x2", mg2, 2.0 * mg)):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err < 1e-5, (name, err)
    nrc.close()


@pytest.mark.gpu
def test_gpu_training_fits_a_fixed_batch(ctx):
    import torch

    rng = np.random.default_rng(2)
    nrc = A.Nrc(A.NrcParams.default(log2_table=16))
    s = make_samples(rng, 8192, valid_frac=1.0)
    p = s["pos"].astype(np.float64) / 64.0  # a smooth target of position
    s["target"] = (1.0 + np.stack([np.sin(3 * p[:, 0]), np.cos(2 * p[:, 1]), p[:, 2]], 1)).astype(np.float32)
    d = torch.from_numpy(s.view(np.uint8).copy()).cuda()
    losses = []
    for step in range(200):
        nrc.train(d.data_ptr(), len(s))
        if step % 50 == 49:
            losses.append(nrc.loss(reset=True))
        else:
            nrc.loss(reset=True)
    print(f"[nrc] losses {losses}", flush=True)
    assert losses[-1] < 0.2 * losses[0] and losses[1] < losses[0]
    nrc.close()
