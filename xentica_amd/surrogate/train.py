"""
Training pairs from the exact engine, training and evaluation.

The exact engine is the ground truth: a training pair is the coarse encoding of
a model's state at step t (16 sub-block means per 16x16 block) and at
step t + horizon, both produced by stepping the model itself.  The
surrogate is scored by the L1 error of its coarse observable (block mean)
against the exact one, next to the persistence baseline (the observable
at step t taken as the prediction), which is what the method gives
without a network.
"""
import numpy as np
import torch

from xentica_amd.core.exceptions import XenticaException
from xentica_amd.surrogate.network import CHANNELS, encode

__all__ = ['coarse_shape', 'value_mask', 'encode_model', 'pairs_from_model',
           'train', 'evaluate']


def coarse_shape(model):
    """``(h, w)`` of the coarse grid of a 2-D model (16x16 blocks)."""
    if len(model.size) != 2 or model.size[0] % 16 or model.size[1] % 16:
        raise XenticaException("surrogate: 2-D fields with sides divisible "
                               "by 16 only")
    return model.size[1] // 16, model.size[0] // 16


def value_mask(model):
    """Byte mask of the first field of a one-byte container."""
    if np.dtype(model.main.dtype).itemsize != 1:
        raise XenticaException("surrogate: one-byte cell states only")
    first = next(iter(model.main.properties.values()))
    return (1 << first.bit_width) - 1


def encode_model(model, x, block_mean=None):
    """Encode the model's current main plane into ``x`` ``[1, h, w, 16]``."""
    model.flush()
    h, w = coarse_shape(model)
    encode(model.gpu.arrays.cells.plane_ptr(model.cells_num), 1, h, w, x,
           block_mean,
           value_mask(model))


def pairs_from_model(model, count, horizon=16, device="cuda"):
    """
    ``count`` consecutive (state at t, state at t + horizon) encodings of
    ``model``, stepping it ``count * horizon`` steps; returns
    ``(inputs, targets)`` of shape ``[count, h, w, 16]`` on the GPU.
    """
    h, w = coarse_shape(model)
    frames = torch.empty((count + 1, h, w, CHANNELS), device=device)
    encode_model(model, frames[0:1])
    for idx in range(count):
        for _ in range(horizon):
            model.step()
        encode_model(model, frames[idx + 1:idx + 2])
    torch.cuda.synchronize()
    return frames[:-1].clone(), frames[1:].clone()


def train(net, inputs, targets, steps=600, batch=4, lr=1e-2, seed=0,
          log=None):
    """
    Adam on the mean squared error of the 16 predicted channels; the
    forward and backward passes are the network's own kernels.  The
    standardisation of the network is set from the inputs' mean and
    standard deviation.  Returns the loss of every step.
    """
    with torch.no_grad():
        net.set_normalization(inputs.mean(), inputs.std())
    gen = torch.Generator().manual_seed(seed)
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    losses = []
    for step in range(steps):
        idx = torch.randint(0, inputs.shape[0], (batch, ), generator=gen)
        idx = idx.to(inputs.device)
        pred = net(inputs[idx])
        loss = torch.mean((pred - targets[idx]) ** 2)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
        losses.append(float(loss.detach()))
        if log and (step % 100 == 0 or step == steps - 1):
            log("surrogate train step %d: mse %.3g" % (step, losses[-1]))
    return losses


def evaluate(net, inputs, targets):
    """
    L1 of the block-mean observable against the exact engine, absolute
    and relative to the mean exact block height, for the surrogate and for
    the two methods without a network: persistence (the observable at step
    t) and the best of the eight one-block rolls of it (the whole coarse
    field moved by one block, the exact answer for a rule that translates
    the field by one block in the horizon).
    """
    with torch.no_grad():
        exact = targets.mean(dim=-1)
        pred = net(inputs).mean(dim=-1)
        persist = inputs.mean(dim=-1)
        scale = float(exact.abs().mean())
        l1_net = float((pred - exact).abs().mean())
        l1_persist = float((persist - exact).abs().mean())
        rolls = {}
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dy or dx:
                    moved = torch.roll(persist, shifts=(dy, dx), dims=(1, 2))
                    rolls[(dy, dx)] = float((moved - exact).abs().mean())
        best = min(rolls, key=rolls.get)
    return {
        "l1": l1_net, "l1_relative": l1_net / scale,
        "persistence_l1": l1_persist,
        "persistence_l1_relative": l1_persist / scale,
        "roll_l1": rolls[best], "roll_l1_relative": rolls[best] / scale,
        "roll_blocks": list(best),
        "mean_block_height": scale,
    }
