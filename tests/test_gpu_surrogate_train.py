"""Training the surrogate of the surrogate-training configuration on the GPU:
a 5-member ensemble of 57 -> 4 x 256 (SiLU) -> 1 MLPs, BF16 MFMA forward and
backward kernels with FP32 accumulation, Adam at lr 1e-3, on a small set the
full FP64 solver generates here (random geometries of 8..100 RyRs on the
31 nm lattice, random open sets, jSR 200..1200 uM, targets ln [Ca2+] at every
mouth).  The loss must fall, and on held-out geometries the ensemble must
beat the superposition estimate it learns the residual of."""
import os

import numpy as np
import pytest

from casim_amd import mlp, surrogate, train_surrogate

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_train_configs_ensemble_lr_1e3(gpu):
    from casim_amd.dyad import DyadHIP
    _, lin = surrogate.load_model(os.path.join(ROOT, "casim_amd", "surrogate_weights.npz"))
    x, y, _ = surrogate.generate(DyadHIP, n_sites=256, seed=300, lin=lin)
    xh, yh, _ = surrogate.generate(DyadHIP, n_sites=96, seed=301, lin=lin)
    print("rows: train %d, held out %d" % (len(y), len(yh)), flush=True)
    assert x.shape[1] == surrogate.N_FEAT + 1 == 57
    yt = y - x[:, surrogate.N_FEAT]  # the residual of the superposition estimate
    net = mlp.MLP(57, width=256, depth=4, members=5, seed=11)
    assert net.P == 57 * 256 + 256 + 3 * (256 * 256 + 256) + 256 + 1
    net.fit_normalization(x, yt)
    net.set_data(x, yt)
    first = net.train(1, batch=512, lr=1e-3, seed=1)
    last = first
    for ep in range(1, 24, 2):
        last = net.train(2, batch=512, lr=1e-3, seed=1 + ep)
    print("loss per member: first epoch", [round(float(v), 4) for v in first], "last", [round(float(v), 4) for v in last],
          flush=True)
    assert np.all(np.isfinite(last))
    assert np.all(last < 0.5 * first)
    held = train_surrogate.errors(net, xh, yh, lin=lin)
    true = np.exp(yh.astype(np.float64))
    sup = float(np.mean(np.abs(np.exp(xh[:, surrogate.N_FEAT].astype(np.float64)) - true) / true))
    print("held-out mean relative error: ensemble %.4f, superposition alone %.4f" % (held["rel_mean"], sup),
          flush=True)
    assert held["rel_mean"] < 0.5 * sup
    assert held["rel_mean"] < 0.05
