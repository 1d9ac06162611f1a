"""
The bit-sliced batch (``runtime/bits_batch.py``, ``csrc/life_batch.hip``)
against the plain engine (the reference's structure run instance by
instance) and the oracle: ragged widths sharing waves, every depth, rule
changes between graph replays, and the 256-instance batch workload at the
launch shapes the autotuner chooses from.
"""
import binascii

import numpy as np
import pytest

from examples.workloads import LifeLikeRule, batch_experiments
from xentica_amd import seeds
from xentica_amd.runtime import hip, kernels
from xentica_amd.runtime.bits_batch import BitsBatch

pytestmark = pytest.mark.gpu


def run(model, steps):
    for _ in range(steps):
        model.step()
    return model


def device_crc(model):
    return kernels.crc32_u8(model.gpu.arrays.cells.device_ptr(),
                            model.cells_num)


def plain_twin(exp, rule=None):
    twin = LifeLikeRule(type("Plain" + exp.__name__, (exp, ),
                             {"engine": "plain"}))
    twin.rule = exp.rule if rule is None else rule
    return twin


def ragged_experiments(sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for idx, size in enumerate(sizes):
        out.append(type("Ragged%d" % idx, (batch_experiments(1)[0], ), {
            "size": size, "word": seed * 1000 + idx,
            "rule": int(rng.integers(0, 1 << 18)),
            "seed": seeds.patterns.PrimordialSoup(
                vals={"state": seeds.random.RandInt(0, 1)})}))
    return out


@pytest.mark.parametrize("depth,strip_rows", [(1, 16), (3, 8), (8, 48),
                                              (5, 24)])
def test_ragged_batch_vs_plain(depth, strip_rows):
    sizes = [(32, 5), (33, 7), (45, 90), (63, 4), (70, 1), (100, 33),
             (1999, 20), (2048, 70), (257, 257), (64, 2)]
    exps = ragged_experiments(sizes, depth)
    models = [LifeLikeRule(exp) for exp in exps]
    batch = BitsBatch(models, depth=depth, strip_rows=strip_rows)
    steps = 2 * depth + 3
    print("ragged bits batch depth %d strips %d" % (depth, strip_rows),
          flush=True)
    batch.run(steps)
    batch.synchronize()
    batch.scatter()
    for exp, model in zip(exps, models):
        twin = run(plain_twin(exp), steps)
        assert device_crc(model) == device_crc(twin), exp.size
        assert np.array_equal(model.gpu.arrays.colors.get(),
                              twin.gpu.arrays.colors.get()), exp.size


def test_rule_changes_and_graph_replays():
    exps = ragged_experiments([(96, 40), (50, 50), (700, 30)], 7)
    models = [LifeLikeRule(exp) for exp in exps]
    eager = BitsBatch([LifeLikeRule(exp) for exp in exps], depth=6)
    batch = BitsBatch(models, depth=6, strip_rows=16)
    rules = [exp.rule for exp in exps]
    rng = np.random.default_rng(3)
    twins = [plain_twin(exp) for exp in exps]
    for round_ in range(4):
        if round_:
            idx = int(rng.integers(0, len(models)))
            rules[idx] = int(rng.integers(0, 1 << 18))
            models[idx].rule = rules[idx]
            eager.models[idx].rule = rules[idx]
        batch.run(20)
        eager.run(20, graph=False)
        for twin, rule in zip(twins, rules):
            twin.rule = rule
            run(twin, 20)
    for item in (batch, eager):
        item.synchronize()
        item.scatter()
    for model, other, twin in zip(models, eager.models, twins):
        assert device_crc(model) == device_crc(twin)
        assert device_crc(other) == device_crc(twin)


@pytest.mark.parametrize("depth,strip_rows", [(8, 16), (8, 48), (4, 24)])
def test_batch_workload_vs_plain_and_oracle(depth, strip_rows):
    """The 256-instance batch, 1000 steps: every instance equal to the
    plain engine run alone with its own rule, the eight smallest also to
    the oracle."""
    from oracle import oracle as O
    exps = batch_experiments()
    models = [LifeLikeRule(exp) for exp in exps]
    starts = [m.gpu.arrays.cells.get()[:m.cells_num] for m in models]
    batch = BitsBatch(models, depth=depth, strip_rows=strip_rows)
    print("bits batch of 256, depth %d, strips %d, 1000 steps"
          % (depth, strip_rows), flush=True)
    batch.run(1000)
    batch.synchronize()
    batch.scatter()
    got = [device_crc(m) for m in models]
    twins = [plain_twin(exp) for exp in exps]
    hip.precompile([m.cuda_source for m in twins],
                   hip.compile_options(LifeLikeRule.fp_contract))
    for idx, (exp, twin) in enumerate(zip(exps, twins)):
        # a parameter is class-wide, as in the reference: set it just
        # before the instance runs
        twin.rule = exp.rule
        run(twin, 1000)
        assert got[idx] == device_crc(twin), "instance %d vs plain" % idx
    for idx in np.argsort([exp.size[0] for exp in exps], kind="stable")[:8]:
        exp = exps[idx]
        field = O.Field(exp.size, 8, np.uint8, starts[idx])
        O.lifelike(field, O.moore_deltas(2), O.moore_rev(2),
                   exp.rule & 0x1FF, (exp.rule >> 9) & 0x1FF, 1000,
                   colors=False)
        assert got[idx] == binascii.crc32(field.cells), \
            "instance %d vs oracle" % idx
