"""CPU pins of the surrogate's float64 reference (oracle/mlp_ref.py) and of the bin quadrature."""
import numpy as np

from oracle import mlp_ref


def small_problem(seed=3, M=6):
    rng = np.random.default_rng(seed)
    params = rng.normal(0, 0.08, mlp_ref.param_count())
    x = rng.uniform(-1, 1, (M, 9))
    t = rng.uniform(0, 1, (M, 9))
    return params, x, t


def test_gradients_match_finite_differences():
    params, x, t = small_problem()
    shift, scale = np.zeros(9), np.ones(9)
    loss, g, _ = mlp_ref.forward_backward(params, x, t, shift, scale)
    rng = np.random.default_rng(0)
    idx = rng.choice(params.size, 40, replace=False)
    for i in idx:
        h = 1e-6
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (mlp_ref.forward_backward(p1, x, t, shift, scale)[0] - mlp_ref.forward_backward(p2, x, t, shift, scale)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_gelu_is_torch_tanh_gelu():
    import torch
    z = np.linspace(-12, 12, 4001)
    ref = torch.nn.functional.gelu(torch.tensor(z, dtype=torch.float64), approximate="tanh").numpy()
    assert np.abs(mlp_ref.gelu(z) - ref).max() < 1e-14
    h = 1e-6
    fd = (mlp_ref.gelu(z + h) - mlp_ref.gelu(z - h)) / (2 * h)
    assert np.abs(mlp_ref.gelu_grad(z) - fd).max() < 1e-8


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 65504.0, 1e-30], np.float32)
    r = mlp_ref.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0           # halfway to even
    assert r[2] == 1.0078125                      # rounds up
    assert abs(r[3] + 3.140625) < 1e-12
    bits = r.astype(np.float32).view(np.uint32)
    assert np.all(bits & 0xFFFF == 0)


def test_two_term_split():
    rng = np.random.default_rng(4)
    a = rng.normal(0, 3, 10000).astype(np.float32)
    tt = mlp_ref.two_term(a)
    hi = mlp_ref.bf16_round(a)
    assert np.all(np.abs(tt - a) <= 2.0 ** -16 * np.abs(a))         # about 16 significant bits
    assert np.all(np.abs(hi - a) <= 2.0 ** -8 * np.abs(a))
    assert np.abs(hi - a).max() > 100 * np.abs(tt - a).max()


def test_two_term_forward_tracks_exact_math():
    """With two-term inputs and activations the bf16 network's outputs stay within 1e-4 of the
    exact network with the same (bf16) weights, where single-term bf16 activations do not."""
    rng = np.random.default_rng(7)
    params = mlp_ref.pack([(mlp_ref.bf16_round(W), b) for W, b in
                           mlp_ref.unpack(rng.normal(0, 0.09, mlp_ref.param_count()))])
    x = rng.uniform(-1, 1, (500, 9))
    t = np.zeros((500, 9))
    shift, scale = np.zeros(9), np.ones(9)
    y_exact = mlp_ref.forward_backward(params, x, t, shift, scale)[2]
    y_bf16 = mlp_ref.forward_backward(params, x, t, shift, scale, bf16=True)[2]
    assert np.abs(y_bf16 - y_exact).max() < 1e-4


def test_gauss_legendre_bin_rule():
    from cudaprob3_amd.surrogate import gl_energies
    E, w = gl_energies([1.0, 10.0])
    assert E.shape == (32,) and abs(w.sum() - 1.0) < 1e-6
    # exact for polynomials up to degree 31 in E over the bin
    x = E[:16]
    for deg in (0, 1, 5, 15, 31):
        exact = ((1.025 ** (deg + 1)) - (0.975 ** (deg + 1))) / ((deg + 1) * 0.05)
        assert abs(float(np.dot(w.astype(np.float64), x ** deg)) - exact) < 1e-6 * max(1.0, exact)
