"""Warm-start MLP: the CPU reference pinned by finite differences (CPU), and the GPU
forward / backward / training kernels checked against it (GPU)."""
import numpy as np
import pytest

from oracle import mlp_ref


def random_params(rng, scale=0.3):
    p = rng.normal(0, scale, mlp_ref.NP)
    p[~mlp_ref.valid_mask()] = 0
    return p


def random_features(rng, n):
    X = np.zeros((n, mlp_ref.INP))
    X[:, :44] = rng.normal(0, 1, (n, 44))
    return X


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = random_params(rng)
    X = random_features(rng, 40)
    Y = rng.normal(0, 1, (40, 3))
    loss, g = mlp_ref.loss_and_grad(p, X, Y, emulated=False)
    idx = np.concatenate([rng.choice(np.nonzero(mlp_ref.valid_mask())[0], 300, replace=False)])
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = mlp_ref.loss_and_grad(q, X, Y, emulated=False)
        q[i] -= 2 * h
        lm, _ = mlp_ref.loss_and_grad(q, X, Y, emulated=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])
    assert np.all(g[~mlp_ref.valid_mask()] == 0)


def test_emulated_close_to_exact():
    rng = np.random.default_rng(1)
    p = random_params(rng)
    X = random_features(rng, 256)
    Y = rng.normal(0, 1, (256, 3))
    l1, g1 = mlp_ref.loss_and_grad(p, X, Y, emulated=False)
    l2, g2 = mlp_ref.loss_and_grad(p, X.astype(np.float16).astype(np.float64), Y, emulated=True)
    assert abs(l1 - l2) < 2e-2 * l1
    assert np.linalg.norm(g1 - g2) < 3e-2 * np.linalg.norm(g1)


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_gpu_forward_matches_reference(ctx):
    import torch
    from ncollide_amd.warmstart import WarmStartMLP
    rng = np.random.default_rng(2)
    mlp = WarmStartMLP(ctx, seed=3)
    p = random_params(rng).astype(np.float32)
    mlp.set_params(p)
    X = random_features(rng, 1000).astype(np.float16)
    out = mlp.forward(torch.as_tensor(X, device=ctx.device)).cpu().numpy()
    ref, _ = mlp_ref.forward(p, X.astype(np.float32), emulated=True)
    err = np.abs(out - ref[:, :3])
    scale = np.abs(ref[:, :3]).max()
    print(f"forward max err {err.max():.3g} (scale {scale:.3g})")
    assert err.max() < 2e-2 * scale
    assert np.mean(err < 1e-4 * scale + 1e-6) > 0.9  # most rows agree to float32 summation order


@pytest.mark.gpu
def test_gpu_gradient_matches_reference(ctx):
    import torch
    from ncollide_amd.warmstart import WarmStartMLP
    rng = np.random.default_rng(4)
    mlp = WarmStartMLP(ctx, seed=5)
    p = random_params(rng, 0.2).astype(np.float32)
    mlp.set_params(p)
    n = 3000
    X = random_features(rng, n).astype(np.float16)
    Y = rng.normal(0, 1, (n, 3)).astype(np.float32)
    g, loss = mlp.grad(torch.as_tensor(X, device=ctx.device), torch.as_tensor(Y, device=ctx.device))
    g = g.cpu().numpy()
    rl, rg = mlp_ref.loss_and_grad(p, X.astype(np.float32), Y, emulated=True)
    assert abs(loss.item() - rl) < 1e-3 * rl
    rel = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    print(f"gradient relative error {rel:.3g}")
    assert rel < 1e-2
    for name, a, b in (("W1", mlp_ref.OFF_W1, mlp_ref.OFF_B1), ("b1", mlp_ref.OFF_B1, mlp_ref.OFF_W2),
                       ("W2", mlp_ref.OFF_W2, mlp_ref.OFF_B2), ("b2", mlp_ref.OFF_B2, mlp_ref.OFF_W3),
                       ("W3", mlp_ref.OFF_W3, mlp_ref.OFF_B3), ("b3", mlp_ref.OFF_B3, mlp_ref.NP)):
        r = np.linalg.norm(g[a:b] - rg[a:b]) / (np.linalg.norm(rg[a:b]) + 1e-12)
        assert r < 2e-2, (name, r)
    assert np.all(g[~mlp_ref.valid_mask()] == 0)


@pytest.mark.gpu
def test_gpu_training_fits_a_smooth_map(ctx):
    import torch
    from ncollide_amd.warmstart import WarmStartMLP
    rng = np.random.default_rng(6)
    n = 65536
    X = random_features(rng, n)
    A = rng.normal(0, 0.3, (44, 3))
    Y = np.tanh(X[:, :44] @ A)
    mlp = WarmStartMLP(ctx, seed=7)
    xt = torch.as_tensor(X.astype(np.float16), device=ctx.device)
    yt = torch.as_tensor(Y.astype(np.float32), device=ctx.device)
    hist = mlp.fit(xt, yt, epochs=6, batch=4096, lr=3e-3)
    print("loss per epoch", hist)
    assert hist[-1] < 0.25 * hist[0]


@pytest.mark.gpu
def test_gpu_descriptors_and_features(ctx):
    import torch
    import ncollide_amd as ncx
    from ncollide_amd import workloads
    from ncollide_amd.warmstart import descriptors, features
    w = workloads.full_pipeline(seed=9, n=400)
    ss = ncx.ShapeSet(records=w["records"], pool=w["pool"], ctx=ctx)
    desc = descriptors(ss).cpu().numpy()
    k = np.arange(16)
    z = 1 - (2 * k + 1) / 16.0
    ph = k * np.pi * (3 - np.sqrt(5))
    U = np.stack([np.sqrt(1 - z * z) * np.cos(ph), np.sqrt(1 - z * z) * np.sin(ph), z], 1)
    rec = w["records"]
    for i in range(400):
        t = rec["type"][i]
        if t == 0:
            ref = np.full(16, rec["p0"][i])
        elif t == 1:
            ref = np.abs(U) @ np.array([rec["p0"][i], rec["p1"][i], rec["p2"][i]])
        elif t == 2:
            ref = rec["p0"][i] * np.abs(U[:, 1]) + rec["p1"][i]
        else:
            pts = w["pool"][rec["hull_offset"][i]:rec["hull_offset"][i] + rec["hull_count"][i]]
            ref = (pts @ U.T).max(0)
        assert np.allclose(desc[i], ref, atol=1e-5), i
    rng = np.random.default_rng(0)
    ia = rng.integers(0, 400, 500)
    ib = rng.integers(0, 400, 500)
    f = features(ss, torch.as_tensor(w["poses"], device=ctx.device), torch.as_tensor(desc, device=ctx.device),
                 ia, ib).cpu().numpy().astype(np.float64)
    P = w["poses"].astype(np.float64)

    def rotm(q):
        x, y, z, w_ = q
        return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w_ * z), 2 * (x * z + w_ * y)],
                         [2 * (x * y + w_ * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w_ * x)],
                         [2 * (x * z - w_ * y), 2 * (y * z + w_ * x), 1 - 2 * (x * x + y * y)]])
    for k in range(500):
        R1, R2 = rotm(P[ia[k], 4:8]), rotm(P[ib[k], 4:8])
        ref = np.concatenate([(R1.T @ R2).ravel(), R1.T @ (P[ib[k], :3] - P[ia[k], :3]), desc[ia[k]], desc[ib[k]],
                              np.zeros(4)])
        assert np.allclose(f[k], ref, atol=2e-3 * np.maximum(1, np.abs(ref)).max()), k
