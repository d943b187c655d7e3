"""Learned start of the sc0 iteration: training data from the solver's own
trajectories, training on the GPU, seeding.

The batch records, per channel, the network's inputs (HF energies and
occupations, the second-order density estimate D2, the first sc0 sweep's
density correction D1, instance descriptors) and the converged correction D
(bcd_batch_record_learning / bcd_batch_learning_data).  The network is
trained by bcd_mlp_fit (BF16 matrix-core forward and backward kernels, Adam)
on the remaining step D - D1 of the channels of converged instances; a batch
given the network (bcd_batch_set_learned_start) starts its second sc0 sweep
from D1 + the prediction, and falls back to the plain iteration for any
instance whose residual exceeds the fallback tolerance."""
from __future__ import annotations

import numpy as np

from . import capi


def collect(batch: "capi.Batch"):
    """Record one solve of `batch` and return (inputs, step targets, instance
    of each row) for the channels of its converged instances; the batch's
    own parameters are used."""
    batch.record_learning(True)
    batch.solve()
    f, t, ci = batch.learning_data()
    batch.record_learning(False)
    conv = np.array([batch.converged(i) for i in range(batch.n)], bool)
    keep = conv[ci] & (f[:, capi.LEARN_NI * 3] > 0)     # converged, channel covered by the network
    x = f[keep]
    y = t[keep].copy()
    y[:, :capi.LEARN_TRI] -= x[:, capi.LEARN_D1:capi.LEARN_D1 + capi.LEARN_TRI]
    return x, y, ci[keep]


def train(x, y, epochs: int = 40, batch: int = 512, lr: float = 1e-3, seed: int = 0, device: int = 0):
    """A network of 3 layers and 256 hidden units trained on the device;
    returns (network, per-epoch mean loss)."""
    net = capi.Mlp(capi.LEARN_IN, 256, capi.LEARN_OUT, seed=seed, device=device)
    hist = net.fit(x, y, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    return net, hist
