"""Surrogate MLP: the reference MLP pinned by finite differences (CPU); the
hand-written MFMA GEMM and the device backward pass against it (GPU)."""
import numpy as np
import pytest

from oracle import mlp_oracle as R


def test_reference_backward_matches_finite_differences():
    rs = np.random.default_rng(0)
    n_in, width, depth = 5, 7, 2
    _, P = R.layout(n_in, width, depth)
    p = rs.normal(0, 0.5, P)
    x = rs.normal(0, 1, (9, n_in))
    y = rs.normal(0, 1, 9)
    loss, g = R.loss_and_grad(p, x, y, n_in, width, depth)
    eps = 1e-6
    for i in rs.choice(P, 25, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (R.loss_and_grad(pp, x, y, n_in, width, depth)[0] - R.loss_and_grad(pm, x, y, n_in, width, depth)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159])
    r = R.to_bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -6
    assert r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12


def _bf16_bits(a):
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(70, 33, 45), (128, 256, 64), (1, 256, 256), (300, 1, 17)])
def test_gemm_bf16_mfma(gpu, ta, tb, M, N, K):
    from casim_amd import mlp as mlpmod
    from tests.helpers import DeviceBuffer
    rs = np.random.default_rng(M * 7 + N * 3 + K + ta * 2 + tb)
    A = rs.normal(0, 1, (K, M) if ta else (M, K))
    B = rs.normal(0, 1, (N, K) if tb else (K, N))
    Ab, Bb = _bf16_bits(A), _bf16_bits(B)
    Cd = DeviceBuffer(np.full((M, N), 0.5, dtype=np.float32))
    da, db = DeviceBuffer(Ab), DeviceBuffer(Bb)
    lib = mlpmod.lib()
    rc = lib.casim_gemm_bf16(da.ptr, A.shape[1], ta, db.ptr, B.shape[1], tb, Cd.ptr, N, M, N, K, 1.0, None)
    assert rc == 0
    Af, Bf = R.to_bf16(A), R.to_bf16(B)
    want = (Af.T if ta else Af) @ (Bf.T if tb else Bf) + 0.5
    got = Cd.read()
    assert np.max(np.abs(got - want)) <= 2e-5 * max(1.0, np.abs(want).max()) * np.sqrt(K)


@pytest.mark.gpu
def test_device_gradient_matches_reference(gpu):
    from casim_amd.mlp import MLP
    rs = np.random.default_rng(3)
    n_in, width, depth = 20, 64, 3
    m = MLP(n_in, width, depth, members=2, seed=5)
    m.set_normalization(np.zeros(n_in), np.ones(n_in), 0.0, 1.0)
    x = rs.normal(0, 1, (200, n_in)).astype(np.float32)
    y = rs.normal(0, 1, 200).astype(np.float32)
    for member in (0, 1):
        p = m.params(member).astype(np.float64)
        loss_d, g_d = m.gradient(member, x, y)
        loss_r, g_r = R.loss_and_grad(p, x.astype(np.float64), y.astype(np.float64), n_in, width, depth, bf16=True)
        assert abs(loss_d - loss_r) <= 1e-3 * abs(loss_r)
        cos = float(g_d @ g_r / np.linalg.norm(g_d) / np.linalg.norm(g_r))
        assert cos > 0.9999
        assert np.linalg.norm(g_d - g_r) <= 5e-3 * np.linalg.norm(g_r)
        # and the unrounded float64 gradient is close too (BF16 rounding only)
        _, g_64 = R.loss_and_grad(p, x.astype(np.float64), y.astype(np.float64), n_in, width, depth, bf16=False)
        assert float(g_d @ g_64 / np.linalg.norm(g_d) / np.linalg.norm(g_64)) > 0.99


@pytest.mark.gpu
def test_training_reduces_loss(gpu):
    from casim_amd.mlp import MLP
    rs = np.random.default_rng(4)
    n_in = 8
    x = rs.uniform(-2, 2, (20000, n_in)).astype(np.float32)
    y = (np.sin(x[:, 0]) + 0.5 * x[:, 1] ** 2 - 0.3 * x[:, 2] * x[:, 3]).astype(np.float32)
    m = MLP(n_in, width=128, depth=3, members=3, seed=9)
    m.fit_normalization(x, y)
    m.set_data(x, y)
    loss1 = m.train(1, batch=1024, lr=1e-3)
    loss2 = m.train(15, batch=1024, lr=1e-3)
    assert np.all(loss2 < 0.05) and np.all(loss2 < loss1 / 5)
    mean, spread, raw = m.predict(x[:512])
    pred = np.log(mean)
    assert np.sqrt(np.mean((raw.mean(axis=0) - y[:512]) ** 2)) < 0.25 * np.std(y)


@pytest.mark.gpu
@pytest.mark.parametrize("n_in,width,depth,members,rows", [(57, 256, 4, 5, 1000), (57, 256, 4, 5, 33),
                                                          (10, 128, 2, 3, 1), (200, 64, 3, 8, 97)])
def test_fused_inference_equals_layer_by_layer(gpu, n_in, width, depth, members, rows):
    """The fused inference kernel (every layer of a member in one launch,
    activations in LDS) gives the per-layer GEMM + bias path's outputs bit for
    bit, for the surrogate's shape, narrower layers, one row and ragged tiles;
    and it follows parameter updates (set_params, training)."""
    from casim_amd import mlp
    gen = np.random.default_rng(n_in + width + rows)
    net = mlp.MLP(n_in, width=width, depth=depth, members=members, seed=5)
    x = gen.normal(0, 2, (rows, n_in)).astype(np.float32)
    net.set_normalization(x.mean(axis=0), x.std(axis=0) + 0.5, 0.3, 1.7)
    for e in range(members):
        net.set_params(e, gen.normal(0, 0.08, net.P).astype(np.float32))

    def both():
        net.set_fused(0)
        a = net.predict(x)
        net.set_fused(1)
        b = net.predict(x)
        return a, b

    a, b = both()
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert np.all(np.isfinite(b[2])) and np.ptp(b[2]) > 0
    # after a parameter update the fused path uses the new weights
    net.set_params(0, gen.normal(0, 0.08, net.P).astype(np.float32))
    a2, b2 = both()
    assert np.array_equal(a2[2], b2[2]) and not np.array_equal(b2[2][0], b[2][0])
    net.set_data(x, gen.normal(0, 1, rows).astype(np.float32))
    net.train(1, batch=min(rows, 64), lr=1e-3, seed=3)
    a3, b3 = both()
    assert np.array_equal(a3[2], b3[2]) and not np.array_equal(b3[2], b2[2])
