"""The surrogate network's matrix-core kernels against the float64 oracle, its binned training
targets against the propagation oracle, and a short training run with the exact fallback."""
import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


def batch(torch, M, seed):
    from cudaprob3_amd.surrogate import FEATURE_HI, FEATURE_LO
    rng = np.random.default_rng(seed)
    x = rng.uniform(FEATURE_LO, FEATURE_HI, (M, 9)).astype(np.float32)
    t = rng.uniform(0, 1, (M, 9)).astype(np.float32)
    return x, t, torch.as_tensor(x, device="cuda"), torch.as_tensor(t, device="cuda")


@pytest.mark.parametrize("kernel", ["layered", "fused4", "fused8"])
def test_forward_matches_oracle(torch, kernel):
    """Two-term activations: the kernels agree with the float64 network (same bf16 weights,
    hi + lo inputs and activations) to FP32 accumulation error."""
    from cudaprob3_amd.surrogate import SurrogateNet
    print(f"[gpu] mlp forward {kernel}", flush=True)
    net = SurrogateNet(max_batch=512, training=False, seed=1)
    net.set_forward_kernel(kernel)
    x, t, xd, _ = batch(torch, 1000, 0)                 # a ragged final tile (and chunk)
    y = net.forward(xd)
    net.sync()
    _, _, y_ref = mlp_ref.forward_backward(net.get_params(), x, t, net.shift, net.scale, bf16=True)
    yg = y.cpu().numpy()
    print(f"  rel {rel(yg, y_ref):.2e} max {np.abs(yg - y_ref).max():.2e}")
    assert rel(yg, y_ref) < 1e-4 and np.abs(yg - y_ref).max() < 2e-4
    net.close()


def test_fused_forward_large_and_empty(torch):
    """Fused inference over many tiles agrees with the layered kernels; n = 0 is a no-op."""
    from cudaprob3_amd.surrogate import SurrogateNet
    print("[gpu] mlp fused vs layered", flush=True)
    net = SurrogateNet(max_batch=8192, training=False, seed=4)
    _, _, xd, _ = batch(torch, 100_003, 5)
    out = {}
    for k in ("layered", "fused4", "fused8"):
        net.set_forward_kernel(k)
        out[k] = net.forward(xd).cpu().numpy()
    assert np.abs(out["fused4"] - out["layered"]).max() < 1e-4
    assert np.abs(out["fused8"] - out["fused4"]).max() < 1e-4
    empty = torch.empty((0, 9), dtype=torch.float32, device="cuda")
    assert net.forward(empty).shape == (0, 9)
    net.close()


def test_gradients_match_oracle(torch):
    from cudaprob3_amd.surrogate import SurrogateNet
    print("[gpu] mlp gradients", flush=True)
    net = SurrogateNet(max_batch=512, training=True, seed=2)
    x, t, xd, td = batch(torch, 512, 1)
    g, loss = net.gradients(xd, td)
    loss_ref, g_ref, _ = mlp_ref.forward_backward(net.get_params(), x, t, net.shift, net.scale, bf16=True)
    assert abs(loss - loss_ref) <= 1e-2 * loss_ref
    off = 0
    for n, k in mlp_ref.SHAPES:
        for size in (n * k, n):
            a, b = g[off:off + size], g_ref[off:off + size]
            assert rel(a, b) < 3e-2, (n, k, size, rel(a, b))
            off += size
    net.close()


def test_bin_average_matches_oracle(oracle_mod):
    from cudaprob3_amd import workloads as W
    from cudaprob3_amd.surrogate import exact_binned, gl_energies
    print("[gpu] GL bin averages", flush=True)
    cos = np.array([-0.95, -0.4, 0.3])
    centers = np.array([0.08, 1.3, 7.0, 400.0])
    p = (0.59, 0.15, 0.8, 2.0, 7.5e-5, -2.45e-3)
    P = exact_binned([(cos, centers, p, 1)])[0].cpu().numpy()        # [9][3][4]
    E, w = gl_energies(centers)
    radii, rhos = W.jitter12_model()
    ref = oracle_mod.propagate(1, cos, E, radii, rhos, 22.0, p).reshape(3, 4, 16, 9)
    ref = np.einsum("cbgt,g->tcb", ref, w.astype(np.float64))
    assert np.abs(P - ref).max() <= 1e-5


def test_training_and_fallback(torch):
    """A short run: the loss falls well below the untrained network's, and after the error-map
    fallback every held-out bin is within the tolerance."""
    from cudaprob3_amd import surrogate as S
    print("[gpu] surrogate short training", flush=True)
    data = S.make_dataset(n_sets=192, n_cos=32, n_bins=32, seed=11)
    held = S.make_dataset(n_sets=32, n_cos=32, n_bins=32, seed=12)
    net = S.SurrogateNet(max_batch=8192, training=True, seed=3)
    y0 = net.forward(held.X)
    net.sync()
    loss0 = float(((y0 - held.T) ** 2).mean())
    S.train(net, data, epochs=12, batch=4096)
    y1 = net.forward(held.X)
    net.sync()
    loss1 = float(((y1 - held.T) ** 2).mean())
    assert loss1 < 0.2 * loss0
    emap = S.build_error_map(net, held, n_cos=8, n_e=8)
    mask = emap.fallback_mask(1e-3)
    err = (y1 - held.T).abs().amax(dim=1).cpu().numpy()
    ic, ie = emap.cell(held.X[:, 0].cpu().numpy(), held.X[:, 1].cpu().numpy())
    kept = ~mask[ic, ie]
    assert kept.sum() == 0 or err[kept].max() <= 1e-3
    net.close()
