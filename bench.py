"""
Headline benchmark: Game of Life B3/S23 on a 16384x16384 torus, through
the reference's API.

One "step" is one generation of the whole field: ``model.step()`` of
``GameOfLife(Life16kExperiment)`` (word 42, ``RandInt(0, 1)`` on every
cell, 50 % density) on the default engine, which runs this model on the
bit-sliced kernel (``xen_life_bits_run``: 64 cells per uint64, 8 steps per
launch in registers).  The state is resident in HBM before the timed
region; the timed region is exactly ``--steps`` calls of ``model.step()``
plus the launch of the last queued steps, between HIP events on the
engine's stream (the null stream), bracketed by device synchronisation.
The final state of the timed run is then compared with the plain engine
(the reference's emit/absorb structure) after the same number of steps
("verified"); a mismatch makes the run fail.

The JSON line also carries:
* ``roofline`` -- algorithmic bytes (one bit read + one bit written per
  cell-update = 0.25 B) over the timed region against MI355X's 8 TB/s HBM
  peak; ``traffic`` is the HBM bytes of one 8-step launch from the
  committed rocprofv3 counters, ``hbm_frac`` that traffic over the timed
  region's average launch time, ``valu_frac`` the committed count of VALU
  instructions per launch over the same time at the VALU issue peak;
* ``cpu_baseline`` -- the same model's generated emit/absorb code compiled
  for the host with OpenMP (``runtime/hostc.py``) on a bounded sample;
* ``plain_engine`` -- the reference-structured GPU path on the same field;
* ``steady_state`` (when ``--steps`` is below 2000) -- 2000 more steps of
  the same field timed the same way: a short run pays each launch's fixed
  cost (about 11 us) and the first launch's wait for ``depth`` queued
  steps on few launches.

Usage: ``python bench.py [--gpus 1] [--steps 2000] [--warmup 200]``.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("cell-updates/s and % HBM roofline; batched and surrogate speedup "
          "vs plain")
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 0.25  # 1 bit read + 1 bit written
# wave64 VALU instructions per second: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4
SIZE = (16384, 16384)
WORD = 42
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "life_bits_16384_traffic.json")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def headline(steps, warmup):
    """
    ``steps`` calls of ``GameOfLife(Life16kExperiment).step()`` after
    ``warmup`` untimed ones, between HIP events on the engine's stream;
    returns ``(model, milliseconds)``.
    """
    from examples.game_of_life import GameOfLife
    from examples.workloads import Life16kExperiment
    from xentica_amd.runtime import hip
    model = GameOfLife(Life16kExperiment)
    engine = model._engine
    if engine.name != "bits":
        raise SystemExit("the default engine of Life 16384^2 is %r, not the "
                         "bit-sliced one" % engine.name)
    # the warm-up runs between the same two events, so that their first
    # record (several microseconds on the host) falls outside the timed run
    start, stop = hip.Event(), hip.Event()
    start.record(None)
    for _ in range(warmup):
        model.step()
    engine.finish()
    stop.record(None)
    stop.synchronize()
    engine.state_ptr  # pylint: disable=pointless-statement
    hip.device().synchronize()
    start.record(None)
    for _ in range(steps):
        model.step()
    engine.finish()
    stop.record(None)
    stop.synchronize()
    hip.device().synchronize()
    return model, start.elapsed_ms(stop)


def steady_state(model, steps=2000):
    """``steps`` more calls of ``model.step()`` (continuing the headline
    field) between HIP events: the rate once the fixed costs of a short
    run (the first launch waiting for ``depth`` queued steps, a shorter
    last launch) are spread over many launches."""
    from xentica_amd.runtime import hip
    engine = model._engine
    engine.finish()
    hip.device().synchronize()
    start, stop = hip.Event(), hip.Event()
    start.record(None)
    for _ in range(steps):
        model.step()
    engine.finish()
    stop.record(None)
    stop.synchronize()
    ms = start.elapsed_ms(stop)
    return {"steps": steps, "ms_per_step": ms / steps,
            "value": model.cells_num * steps / (ms / 1e3),
            "unit": "cell-updates/s"}


def counters():
    """The committed rocprofv3 counter summary of the headline kernel."""
    if not os.path.exists(TRAFFIC_FILE):
        return {}
    with open(TRAFFIC_FILE) as summary:
        return json.load(summary)


def host_baseline(model, steps, sample, warm=1):
    """
    The model's generated emit/absorb code compiled for the host with
    OpenMP (``runtime/hostc.py``: the reference's plain structure, colours
    included) on this machine's cores: ``steps`` steps after ``warm``
    untimed ones, from the model's step-0 state.
    """
    from xentica_amd.runtime.hostc import PlainHost, host_threads
    host = PlainHost(model)
    host.load(model.gpu.arrays.cells.get(), model.gpu.arrays.colors.get())
    cores = host_threads()
    host.step(warm, threads=cores)
    t0 = time.perf_counter()
    host.step(steps, threads=cores)
    elapsed = time.perf_counter() - t0
    return {"value": model.cells_num * steps / elapsed,
            "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": sample}


def cpu_baseline(steps=4):
    """The headline model's generated code on the host cores, ``steps``
    steps of the same 16384^2 field."""
    from examples.game_of_life import GameOfLife
    from examples.workloads import Life16kExperiment
    cpu_exp = type("Life16kHost", (Life16kExperiment, ), {"engine": "plain"})
    return host_baseline(
        GameOfLife(cpu_exp), steps,
        "GameOfLife 16384x16384 word 42, %d steps after 1 warm-up step: the "
        "DSL's generated emit/absorb (8 buffer planes, MovingAverage "
        "colours) compiled for the host, g++ -O3 -fopenmp" % steps)


def plain_reference(total_steps, timed_steps=20):
    """
    The plain engine (the reference's emit/absorb structure) on the
    headline field: the device CRC after ``total_steps`` steps, and its
    cell-updates/s over ``timed_steps`` of them.
    """
    from examples.game_of_life import GameOfLife
    from examples.workloads import Life16kExperiment
    from xentica_amd.runtime import hip, kernels
    plain_exp = type("PlainLife16k", (Life16kExperiment, ),
                     {"engine": "plain"})
    model = GameOfLife(plain_exp)
    timed_steps = max(0, min(timed_steps, total_steps - 1))
    rest = total_steps - timed_steps
    for _ in range(rest):
        model.step()
    hip.device().synchronize()
    t0 = time.perf_counter()
    for _ in range(timed_steps):
        model.step()
    hip.device().synchronize()
    elapsed = time.perf_counter() - t0
    crc = kernels.crc32_u8(model.gpu.arrays.cells.plane_ptr(model.cells_num),
                           model.cells_num)
    rate = {"value": model.cells_num * timed_steps / elapsed,
            "unit": "cell-updates/s", "steps": timed_steps,
            "structure": "generated emit + absorb, 8 buffer planes + "
                         "colours in HBM, one launch pair per step"} \
        if timed_steps else None
    return crc, rate


def life512(steps=1000, cpu=True):
    """
    Life B3/S23 512x512 (word 42, 50 %) through the drop-in API, where the
    per-step host work of ``model.step()`` dominates: ``steps`` steps on
    the default engine (bits: launches deferred until a launch is full),
    the same with the reference's test idiom (``crc32`` of
    ``cells.get()`` after every step), the fused and the plain engine; the
    every-step hashes are checked against the oracle's.
    """
    import binascii
    from examples.game_of_life import GameOfLife
    from examples.workloads import Life512Experiment
    from xentica_amd.runtime import hip
    n = 512 * 512

    def timed_run(engine, hashed):
        exp = type("Life512" + engine.title(), (Life512Experiment, ),
                   {"engine": engine})
        model = GameOfLife(exp)
        # one step before the clock starts: kernels compile, load and
        # launch for the first time, work buffers are allocated
        model.step()
        model.gpu.arrays.cells.get()
        hip.device().synchronize()
        hashes = []
        t0 = time.perf_counter()
        for _ in range(steps):
            model.step()
            if hashed:
                hashes.append(binascii.crc32(
                    model.gpu.arrays.cells.get()[:model.cells_num]))
        model.gpu.arrays.cells.get()
        hip.device().synchronize()
        return time.perf_counter() - t0, hashes

    def rate(seconds):
        return {"value": n * steps / seconds, "unit": "cell-updates/s",
                "us_per_step": seconds / steps * 1e6}

    bits_s, _ = timed_run("bits", False)
    hashed_s, hashes = timed_run("bits", True)
    fused_s, _ = timed_run("fused", False)
    plain_s, _ = timed_run("plain", False)
    out = {
        "workload": "Life B3/S23 512x512 torus, word 42, RandInt(0, 1), "
                    "%d steps through model.step() after one warm-up step"
                    % steps,
        "bits": rate(bits_s), "bits_crc_every_step": rate(hashed_s),
        "fused": rate(fused_s), "plain": rate(plain_s),
    }
    out["crc_step_%d" % (steps + 1)] = hashes[-1]
    if cpu:
        # the checker of the every-step hashes runs with the CPU baseline
        out["hashes_match_oracle"] = life512_oracle_hashes(steps) == hashes
        cpu_exp = type("Life512Host", (Life512Experiment, ),
                       {"engine": "plain"})
        out["cpu_baseline"] = host_baseline(
            GameOfLife(cpu_exp), steps,
            "the whole workload: %d steps after 1 warm-up step, the DSL's "
            "generated emit/absorb (MovingAverage colours) compiled for the "
            "host, g++ -O3 -fopenmp" % steps)
    return out


def life512_oracle_hashes(steps):
    """The oracle's CRC of steps 2 .. steps + 1 of Life 512^2 (the checker
    of the timed runs, which start after one warm-up step)."""
    from examples.game_of_life import GameOfLife
    from examples.workloads import Life512Experiment
    from oracle import oracle
    n = 512 * 512
    field = oracle.Field((512, 512), 8, np.uint8,
                         GameOfLife(Life512Experiment).gpu.arrays.cells.get()
                         [:n])
    want = oracle.lifelike(field, oracle.moore_deltas(2), oracle.moore_rev(2),
                           8, 12, steps + 1, hashes=True)
    return [int(h) for h in want[1:]]


def render_leg(steps=7, frames=5, launches=20):
    """
    Frames of a 4096x4096 Life field (viewport = field, zoom 1): ``steps``
    calls of ``model.step()``, then ``render()``.  Engines: bits (the
    packed state projected straight into the image), fused with the render
    epilogue (the queued steps' last launch writes the image from the
    colours in LDS) and fused with the steps flushed first (colours to the
    int3 array, then the render kernel).  Each frame's image is checked
    against the plain engine's; timed: the whole call (kernels, device cast
    to uint8, copy to the host) and the device part alone (steps plus image
    kernels, HIP events).
    """
    from examples.game_of_life import GameOfLife
    from examples.workloads import Life512Experiment
    from xentica_amd.runtime import hip
    size = (4096, 4096)
    pixels = size[0] * size[1]
    models = {}
    for name, engine in (("bits", "bits"), ("fused_epilogue", "fused"),
                         ("fused_flush_then_render", "fused"),
                         ("plain", "plain")):
        exp = type("Render4k" + engine.title(), (Life512Experiment, ),
                   {"size": size, "engine": engine})
        model = GameOfLife(exp)
        model.set_viewport(size)
        models[name] = model

    def frame(name):
        model = models[name]
        for _ in range(steps):
            model.step()
        if name == "fused_flush_then_render":
            model._engine.flush()
        return model.render()

    def device_part(name):
        model = models[name]
        params = [param.dtype(param.value) for param in model._render_params]
        for _ in range(steps):
            model.step()
        if name == "fused_flush_then_render":
            model._engine.flush()
            args = (model.renderer.get_args_vals(model) + params
                    + [np.int64(pixels)])
            model._engine.launch_elementwise(model.gpu.kernels.render, args,
                                             pixels)
        else:
            model._engine.render_frame(params, pixels)

    out = {"workload": "frames of Life 4096x4096: %d model.step() calls, "
                       "then render(); viewport 4096x4096, zoom 1" % steps}
    names = ("bits", "fused_epilogue", "fused_flush_then_render")
    for name in names:
        out[name] = {"images_match_plain_engine": True}
    for _ in range(2):
        want = frame("plain")
        for name in names:
            if not np.array_equal(frame(name), want):
                out[name]["images_match_plain_engine"] = False
    out["fused_epilogue"]["epilogue_frames"] = \
        models["fused_epilogue"]._engine.epilogue_frames
    for name in names:
        hip.device().synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            frame(name)
        out[name]["frame_ms"] = (time.perf_counter() - t0) / frames * 1e3
        start, stop = hip.Event(), hip.Event()
        device_part(name)
        hip.device().synchronize()
        start.record(None)
        for _ in range(launches):
            device_part(name)
        stop.record(None)
        stop.synchronize()
        out[name]["device_us_per_frame"] = \
            start.elapsed_ms(stop) / launches * 1e3
    return out


def batched(steps=1000, plain_steps=10, cpu=True, depth=0, strip_rows=0):
    """
    The batch workload: 256 Life-like tori, sides 64-2048, own rules.

    Bit-sliced batch: one launch per ``depth`` steps over every instance
    (depth and strip height chosen by the autotuner), the run captured in
    a HIP graph.  Checked after the run: every instance's state equals the
    fused LDS-tile batch's (an independent kernel) after the same steps.
    Baselines: the fused batch, and the per-instance plain launch loop (two
    launches per instance per step) over ``plain_steps`` steps.
    """
    from examples.workloads import LifeLikeRule, batch_experiments
    from xentica_amd.runtime import hip, kernels
    from xentica_amd.runtime.bits_batch import autotune as bits_autotune
    from xentica_amd.runtime.fused import Batch
    exps = batch_experiments()
    models = [LifeLikeRule(exp) for exp in exps]
    total = sum(model.cells_num for model in models)
    if depth:
        from xentica_amd.runtime.bits_batch import BitsBatch
        batch, tuned = BitsBatch(models, depth=depth,
                                 strip_rows=strip_rows or 48), {}
    else:
        batch, tuned = bits_autotune(models, log=log)

    def timed_run(item):
        item.run(steps)                       # capture, warm up
        item.synchronize()
        start, stop = hip.Event(), hip.Event()
        start.record(item.stream)
        item.run(steps)
        stop.record(item.stream)
        stop.synchronize()
        return start.elapsed_ms(stop)

    ms = timed_run(batch)
    value = total * steps / (ms / 1e3)
    batch.scatter()
    bits_crcs = [kernels.crc32_u8(m.gpu.arrays.cells.plane_ptr(m.cells_num),
                                  m.cells_num) for m in models]
    fused_models = [LifeLikeRule(exp) for exp in exps]
    fused = Batch(fused_models, tile=(64, 32), depth=8)
    fused_ms = timed_run(fused)
    fused_value = total * steps / (fused_ms / 1e3)
    fused.scatter()
    same = bits_crcs == [kernels.crc32_u8(m.gpu.arrays.cells.plane_ptr(m.cells_num),
                                          m.cells_num) for m in fused_models]
    plain_value = None
    if plain_steps:
        plain_models = []
        for exp in exps:
            plain_exp = type(exp.__name__ + "Plain", (exp,),
                             {"engine": "plain"})
            plain_models.append(LifeLikeRule(plain_exp))
        hip.precompile([model.cuda_source for model in plain_models],
                       hip.compile_options(LifeLikeRule.fp_contract))
        for model, exp in zip(plain_models, exps):
            model.rule = exp.rule
            model.step()
        hip.device().synchronize()
        t0 = time.perf_counter()
        for _ in range(plain_steps):
            for model, exp in zip(plain_models, exps):
                model.rule = exp.rule
                model.step()
        hip.device().synchronize()
        plain_value = total * plain_steps / (time.perf_counter() - t0)
    out = {
        "workload": "256 Life-like tori, sides log-uniform 64-2048, random "
                    "9-bit birth/survival masks, density 0.1-0.6, %d steps"
                    % steps,
        "cells": total, "value": value, "unit": "cell-updates/s",
        "ms": ms, "engine": "bit-sliced batch, depth %d, strips of %d rows, "
                            "%d waves, graph" % (batch.depth,
                                                 batch.strip_rows,
                                                 batch.n_waves),
        "lane_use": batch.lane_use(),
        "algorithmic_bytes_per_cell_update": 0.25,
        "algorithmic_GBps": value * 0.25 / 1e9,
        "roofline_frac": value * 0.25 / 1e9 / HBM_PEAK_GBS,
        "uint8_layout_GBps": value * 2.0 / 1e9,
        "autotune": {"d%d/r%d" % key: rate for key, rate in tuned.items()},
        "instances_equal_fused_batch": same,
        "fused_batch": {"value": fused_value, "unit": "cell-updates/s",
                        "ms": fused_ms, "tile": [64, 32], "depth": 8},
        "plain_loop_value": plain_value, "plain_loop_steps": plain_steps,
        "speedup_vs_plain": value / plain_value if plain_value else None,
        "speedup_vs_fused_batch": value / fused_value,
    }
    if cpu:
        out["cpu_baseline"] = batch_cpu_baseline(exps)
    return out


def batch_cpu_baseline(exps, sample=16, steps=10):
    """Every 16th instance by side, each model's generated code on the
    host cores for ``steps`` steps (the rate over the sample's cells)."""
    from examples.workloads import LifeLikeRule
    from xentica_amd.runtime.hostc import PlainHost, host_threads
    order = sorted(range(len(exps)), key=lambda idx: exps[idx].size[0])
    picks = order[::max(1, len(order) // sample)][:sample]
    cores = host_threads()
    cells = 0
    elapsed = 0.0
    for idx in picks:
        exp = exps[idx]
        model = LifeLikeRule(type("Host%d" % idx, (exp, ),
                                  {"engine": "plain"}))
        model.rule = exp.rule
        host = PlainHost(model)
        host.load(model.gpu.arrays.cells.get(), model.gpu.arrays.colors.get())
        host.step(1, threads=cores)
        t0 = time.perf_counter()
        host.step(steps, threads=cores)
        elapsed += time.perf_counter() - t0
        cells += model.cells_num
    return {"value": cells * steps / elapsed, "unit": "cell-updates/s",
            "cores": cores, "kind": "port",
            "sample": "%d of the 256 instances (every 16th by side, %d "
                      "cells), %d steps each after 1 warm-up step: the "
                      "generated emit/absorb code compiled for the host "
                      "(g++ -O3 -fopenmp), one instance after another"
                      % (len(picks), cells, steps)}


def life3d(steps=200, cpu=True):
    """3-D B5/S45, 1024^3 torus, 10 % density: packed vs unpacked engines."""
    from examples.workloads import life3d_cells
    from xentica_amd.runtime import hip
    from xentica_amd.runtime.lifebits import Life3DField, Life3DBytes
    size = (1024, 1024, 1024)
    n = size[0] * size[1] * size[2]
    born, sustain = 1 << 5, (1 << 4) | (1 << 5)
    cells = life3d_cells(size)
    stream = hip.Stream()
    out = {"workload": "3-D Life-like B5/S45, reference 26-offset Moore list, "
                       "1024^3 torus, 10 %% density, %d steps; uint8: one "
                       "byte per cell in and out, packed once and expanded "
                       "once inside the timed region" % steps}
    legs = (("bitpacked", Life3DField, {}, 0.25, steps),
            ("uint8", Life3DBytes, {}, 2.0, steps),
            ("uint8_direct", Life3DBytes, {"direct": True}, 2.0,
             min(20, steps)))
    for name, cls, kwargs, bpc, count in legs:
        field = cls(size, born, sustain, stream=stream, **kwargs)
        field.set_cells(cells)
        field.step(2)
        # the uint8 plane is packed by the first step and expanded by the
        # read at the end: both inside the timed region
        field.state_ptr  # pylint: disable=pointless-statement
        start, stop = hip.Event(), hip.Event()
        stream.synchronize()
        start.record(stream)
        field.step(count)
        field.state_ptr  # pylint: disable=pointless-statement
        stop.record(stream)
        stop.synchronize()
        ms = start.elapsed_ms(stop)
        value = n * count / (ms / 1e3)
        out[name] = {"value": value, "unit": "cell-updates/s",
                     "ms_per_step": ms / count, "steps": count,
                     "algorithmic_GBps": value * bpc / 1e9,
                     "roofline_frac": value * bpc / 1e9 / HBM_PEAK_GBS}
        if name == "bitpacked":
            crc_want = field.crc32()
        del field
    out["dsl"] = life3d_dsl(steps, crc_want)
    if cpu:
        out["cpu_baseline"] = life3d_cpu_baseline(born, sustain)
    return out


def life3d_dsl(steps, crc_want):
    """``LifeLike3D(Life3DExperiment).step()`` at 1024^3 on the default
    engine: ``steps`` steps after 2, between HIP events; the final state is
    compared with the packed kernel's run of the same cells."""
    from examples.workloads import LifeLike3D, Life3DExperiment
    from xentica_amd.runtime import hip
    t0 = time.perf_counter()
    model = LifeLike3D(Life3DExperiment)
    build_s = time.perf_counter() - t0
    engine = model._engine
    for _ in range(2):
        model.step()
    engine.finish()
    hip.device().synchronize()
    start, stop = hip.Event(), hip.Event()
    start.record(None)
    for _ in range(steps):
        model.step()
    engine.finish()
    stop.record(None)
    stop.synchronize()
    ms = start.elapsed_ms(stop)
    value = model.cells_num * steps / (ms / 1e3)
    return {"api": "LifeLike3D(Life3DExperiment).step()",
            "engine": engine.name, "value": value, "unit": "cell-updates/s",
            "ms_per_step": ms / steps, "steps": steps,
            "model_build_and_seed_s": build_s,
            "state_equals_packed_kernel": engine.crc32() == crc_want}


def life3d_cpu_baseline(born, sustain, side=256, steps=10):
    """``LifeLike3D``'s generated code on the host cores on a 256^3 sample
    of the same workload (the reference's plain layout of 1024^3 needs 27
    GiB of buffer planes): same rule, 26-offset list and 10 % seeding."""
    from examples.workloads import LifeLike3D, Life3DExperiment
    cpu_exp = type("Life3DHost", (Life3DExperiment, ),
                   {"size": (side, side, side), "engine": "plain",
                    "born": born, "sustain": sustain})
    return host_baseline(
        LifeLike3D(cpu_exp), steps,
        "%d^3 torus (the 1024^3 field in the reference's plain layout needs "
        "27 GiB), B5/S45 with the 26-offset list, 10 %% density, %d steps "
        "after 1 warm-up step: the DSL's generated emit/absorb (26 buffer "
        "planes, colours) compiled for the host, g++ -O3 -fopenmp"
        % (side, steps))


def energy_cpu_baseline(steps=10):
    """EnergyLife's generated code on the host cores, the same 4096^2
    field, ``steps`` steps (colours included, as the GPU computes them)."""
    from examples.energy_life import EnergyLife, EnergyLifeExperiment
    cpu_exp = type("EnergyHost", (EnergyLifeExperiment, ),
                   {"engine": "plain"})
    return host_baseline(
        EnergyLife(cpu_exp), steps,
        "EnergyLife 4096x4096 word 42, %d steps after 1 warm-up step: the "
        "DSL's generated emit/absorb (8 buffer planes of 64-bit words, "
        "colours) compiled for the host, g++ -O3 -fopenmp" % steps)


def energy_life(steps=5000, plain_steps=20, cpu=True, tiles=None,
                depths=None):
    """
    EnergyLife on a 4096x4096 torus: 64-bit words (energy 8 bits, genome
    18, mode 1, xorshift32 32), word 42, ``steps`` steps by the fused
    engine in one graph-captured run; total energy checked before and
    after; the plain emit/absorb launch loop timed over ``plain_steps``.
    """
    from examples.energy_life import (
        EnergyLife, EnergyLifeExperiment, field_values)
    from xentica_amd.runtime import hip
    from xentica_amd.runtime.fused import autotune
    model = EnergyLife(EnergyLifeExperiment)
    n = model.cells_num
    total0 = int(field_values(model, 'energy').sum())
    batch, _ = autotune([model], log=log, **_tune_space(tiles, depths))
    batch.run(10)
    batch.synchronize()
    start, stop = hip.Event(), hip.Event()
    start.record(batch.stream)
    batch.run(steps)
    stop.record(batch.stream)
    stop.synchronize()
    ms = start.elapsed_ms(stop)
    batch.scatter()
    energy = field_values(model, 'energy')
    mode = field_values(model, 'mode')
    value = n * steps / (ms / 1e3)
    plain_value = None
    if plain_steps:
        plain_exp = type("EnergyLifePlain", (EnergyLifeExperiment, ),
                         {"engine": "plain"})
        plain = EnergyLife(plain_exp)
        plain.step()
        hip.device().synchronize()
        t0 = time.perf_counter()
        for _ in range(plain_steps):
            plain.step()
        hip.device().synchronize()
        plain_value = n * plain_steps / (time.perf_counter() - t0)
    bpc = 16.0
    return {
        "workload": "EnergyLife 4096x4096 torus, 64-bit words (energy 8, "
                    "genome 18, mode 1, xorshift32), word 42, %d steps"
                    % steps,
        "value": value, "unit": "cell-updates/s", "ms_per_step": ms / steps,
        "engine": "fused, tile %s depth %d, graph" % (
            list(batch.program.tile), batch.program.depth),
        "algorithmic_bytes_per_cell_update": bpc,
        "algorithmic_GBps": value * bpc / 1e9,
        "roofline_frac": value * bpc / 1e9 / HBM_PEAK_GBS,
        "energy_total_start": total0, "energy_total_end": int(energy.sum()),
        "energy_conserved": int(energy.sum()) == total0,
        "alive_fraction_end": float(mode.mean()),
        "plain_loop_value": plain_value, "plain_loop_steps": plain_steps,
        "speedup_vs_plain": value / plain_value if plain_value else None,
        **({"cpu_baseline": energy_cpu_baseline()} if cpu else {}),
    }


def _tune_space(tiles, depths):
    """Keyword arguments of ``autotune`` restricting its candidates."""
    space = {}
    if tiles:
        space["tiles"] = tiles
    if depths:
        space["depths"] = depths
    return space


def _event_ms(stream, work, repeats):
    """Average milliseconds of ``work()`` between HIP events on ``stream``."""
    from xentica_amd.runtime import hip
    work()
    start, stop = hip.Event(), hip.Event()
    stream.synchronize()
    start.record(stream)
    for _ in range(repeats):
        work()
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_ms(stop) / repeats


def surrogate(train_words=(101, 102, 103, 104, 105, 106), pairs=3,
              eval_word=199, train_steps=1500, repeats=20, cpu=True):
    """
    The neural-CA surrogate on 4096x4096 sliding sand (heights 0-8, word
    42 family; ``examples/sand_slide.py``: half a block of drift in 16
    steps and slides down steep edges): trained on exact-engine pairs
    (state at t, state at t + 16), scored on a held-out field against the
    exact engine, persistence and the best one-block roll, and timed to the
    coarse field (256x256 block means 16 steps ahead) against the exact
    engine.
    """
    import torch
    from examples.sand_slide import SandSlide, slide_experiment
    from xentica_amd.runtime.fused import autotune
    from xentica_amd.surrogate.network import (
        NCASurrogate, NCAFunction, encode, forward_raw, predict)
    from xentica_amd.surrogate.train import (
        pairs_from_model, train, evaluate, value_mask)
    size = (4096, 4096)
    h, w = size[1] // 16, size[0] // 16
    data = [pairs_from_model(SandSlide(slide_experiment(word, size)), pairs)
            for word in train_words]
    inputs = torch.cat([d[0] for d in data])
    targets = torch.cat([d[1] for d in data])
    eval_in, eval_out = pairs_from_model(
        SandSlide(slide_experiment(eval_word, size)), 2)
    net = NCASurrogate(seed=0)
    t0 = time.perf_counter()
    losses = train(net, inputs, targets, steps=train_steps, batch=4)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    score = evaluate(net, eval_in, eval_out)

    # time to the coarse field: exact engine (fused, autotuned, graph) for
    # 16 steps + block means, against encoder + surrogate forward
    model = SandSlide(slide_experiment(eval_word, size))
    mask = value_mask(model)
    batch, _ = autotune([model], trial_steps=16)
    stream = batch.stream
    x = torch.empty((1, h, w, 16), device="cuda")
    y = torch.empty_like(x)
    mean = torch.empty((1, h, w), device="cuda")
    pred = torch.empty((1, h, w), device="cuda")
    params = net.inference_params()
    torch.cuda.synchronize()

    def exact():
        batch.run(16)
        encode(batch.state[batch.cur].ptr, 1, h, w, x, mean, mask, stream)

    def two_launches():
        encode(batch.state[batch.cur].ptr, 1, h, w, x, None, mask, stream)
        forward_raw(x, params, y, pred, stream)

    def fast():
        predict(batch.state[batch.cur].ptr, 1, h, w, params, None, pred, mask,
                stream)

    exact_ms = _event_ms(stream, exact, repeats)
    two_ms = _event_ms(stream, two_launches, repeats)
    fast_ms = _event_ms(stream, fast, repeats)

    # the reference-structured plain engine for the same 16 steps
    plain = SandSlide(type("SandPlain", (slide_experiment(eval_word, size),),
                           {"engine": "plain"}))
    plain.step()
    hip_sync = __import__("xentica_amd.runtime.hip", fromlist=["hip"])
    hip_sync.device().synchronize()
    t0 = time.perf_counter()
    for _ in range(16):
        plain.step()
    hip_sync.device().synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3

    # kernel throughput at a training-sized batch
    big = inputs[:16].contiguous()
    big_y = torch.empty_like(big)
    cells = big.shape[0] * h * w
    ktensors = [t.detach().clone().requires_grad_()
                for t in net.kernel_tensors()]
    fwd_ms = _event_ms(stream, lambda: forward_raw(big, params, big_y, None,
                                                   stream), repeats)
    grad_out = torch.ones_like(big)

    def fwd_bwd():
        NCAFunction.apply(big, *ktensors).backward(grad_out)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        fwd_bwd()
    torch.cuda.synchronize()
    fb_ms = (time.perf_counter() - t0) * 1e3 / repeats
    mfma_flop = 2 * (16 * 64 + 64 * 16)
    bytes_cell = 16 * 4 * 2
    return {
        "workload": "neural-CA surrogate on 4096x4096 sliding sand "
                    "(heights 0-8, half a block of drift per 16 steps, "
                    "slides down steep edges): 3x3 depthwise (16 ch) + 1x1 "
                    "MLP 16-64-16, bf16 MFMA / fp32 accumulate; predicts "
                    "256x256 block means 16 steps ahead",
        "train": {"pairs": int(inputs.shape[0]), "steps": train_steps,
                  "seconds": train_s, "loss_first": losses[0],
                  "loss_last": losses[-1]},
        "heldout": score,
        "time_to_coarse_field_ms": {"surrogate": fast_ms,
                                    "surrogate_two_launches": two_ms,
                                    "exact_fused": exact_ms,
                                    "exact_plain": plain_ms},
        "surrogate_path": "xen_nca_predict: fine field to block means in one "
                          "launch (encode into LDS, then the forward)",
        "speedup_vs_exact": exact_ms / fast_ms,
        "speedup_vs_plain": plain_ms / fast_ms,
        "exact_engine": "fused, tile %s depth %d, graph" % (
            list(batch.program.tile), batch.program.depth),
        "forward_kernel": {
            "cells": cells, "ms": fwd_ms,
            "mfma_TFLOPs": mfma_flop * cells / (fwd_ms / 1e3) / 1e12,
            "mfma_frac_of_bf16_peak":
                mfma_flop * cells / (fwd_ms / 1e3) / 2.5e15,
            "algorithmic_GBps": bytes_cell * cells / (fwd_ms / 1e3) / 1e9,
            "hbm_frac": bytes_cell * cells / (fwd_ms / 1e3) / 1e9
                        / HBM_PEAK_GBS},
        "forward_backward_ms": fb_ms,
        **({"cpu_baseline": surrogate_cpu_baseline(eval_word, size, mask)}
           if cpu else {}),
    }


def surrogate_cpu_baseline(word, size, mask, horizon=16):
    """
    Time to the coarse field on the host: the sliding-sand model's
    generated emit/absorb compiled for the host cores runs ``horizon``
    steps of the held-out field, then NumPy takes the 16x16 block means
    of the heights.
    """
    from examples.sand_slide import SandSlide, slide_experiment
    from xentica_amd.runtime.hostc import PlainHost, host_threads
    model = SandSlide(type("SandHost", (slide_experiment(word, size), ),
                           {"engine": "plain"}))
    host = PlainHost(model)
    host.load(model.gpu.arrays.cells.get(), model.gpu.arrays.colors.get())
    cores = host_threads()
    host.step(1, threads=cores)  # warm the pages and the thread pool
    host.load(model.gpu.arrays.cells.get(), model.gpu.arrays.colors.get())
    t0 = time.perf_counter()
    host.step(horizon, threads=cores)
    grains = (host.cells & mask).reshape(size[1] // 16, 16, size[0] // 16,
                                         16)
    means = grains.mean(axis=(1, 3), dtype=np.float32)
    elapsed = time.perf_counter() - t0
    return {"value": elapsed * 1e3, "unit": "ms to the coarse field",
            "cores": cores, "kind": "port",
            "cell_updates_per_s": model.cells_num * horizon / elapsed,
            "coarse_mean": float(means.mean()),
            "sample": "the whole held-out workload: %d steps of 4096x4096 "
                      "sliding sand (the DSL's generated emit/absorb, 8 "
                      "buffer planes, g++ -O3 -fopenmp) and the 256x256 "
                      "block means" % horizon}


def refbench(steps=100000, plain_steps=10000, cpu=True, cpu_steps=200):
    """
    The reference's own benchmark (``tests/benchmark.py``: Conway's Life
    640x360, Shifting Sands, Noise TV and EvoLife through ``model.step()``,
    ``NUM_STEPS * cells_num / seconds``), on the default engine with the
    reference's 100000 steps and on the plain engine with fewer.
    """
    from examples import benchmark
    out = {"workload": "the reference's tests/benchmark.py models through "
                       "model.step(); %d steps (plain engine %d)"
                       % (steps, plain_steps)}
    for engine, count in ((None, steps), ("plain", plain_steps)):
        if not count:
            continue
        for name, (speed, cells, used) in benchmark.run(
                count, engine, log=log).items():
            out.setdefault(name, {"cells": cells})[used] = {
                "value": speed, "unit": "cells/s",
                "us_per_step": cells / speed * 1e6, "steps": count}
    if cpu:
        for name, model, experiment in benchmark.MODELS:
            host_exp = type(experiment.__name__ + "Host", (experiment, ),
                            {"engine": "plain"})
            out[name]["cpu_baseline"] = host_baseline(
                model(host_exp), cpu_steps,
                "%d steps after 1 warm-up step: the model's generated "
                "emit/absorb compiled for the host, g++ -O3 -fopenmp"
                % cpu_steps)
    return out


_LEGS = {
    "life512": life512,
    "refbench": refbench,
    "render": render_leg,
    "life3d": life3d,
    "energy": energy_life,
    "batch": batched,
    "surrogate": surrogate,
}


def main():
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=2000)
    parser.add_argument("--warmup", type=int, default=200)
    parser.add_argument("--no-cpu", action="store_true")
    parser.add_argument("--no-plain", action="store_true")
    parser.add_argument("--no-batch", action="store_true")
    parser.add_argument("--no-3d", action="store_true")
    parser.add_argument("--no-energy", action="store_true")
    parser.add_argument("--no-surrogate", action="store_true")
    parser.add_argument("--no-512", action="store_true")
    parser.add_argument("--no-render", action="store_true")
    parser.add_argument("--no-refbench", action="store_true")
    parser.add_argument("--leg-steps", type=int, default=0,
                        help="steps of the workload with --only life3d, "
                             "energy, batch or life512 (0 = the leg's "
                             "default)")
    parser.add_argument("--fused-tile", default="",
                        help="WxH: the fused legs (batch, energy) use this "
                             "tile instead of autotuning it")
    parser.add_argument("--fused-depth", type=int, default=0,
                        help="the fused legs use this depth instead of "
                             "autotuning it")
    parser.add_argument("--bits-depth", type=int, default=0,
                        help="the bit-sliced batch runs this depth instead "
                             "of autotuning (with --only batch)")
    parser.add_argument("--bits-strips", type=int, default=0,
                        help="strip rows of the bit-sliced batch with "
                             "--bits-depth (0: 48)")
    parser.add_argument("--only", choices=sorted(_LEGS),
                        help="run one secondary workload alone (profiling)")
    args = parser.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        return 2
    if args.only:
        log("%s only ..." % args.only)
        kwargs = ({"cpu": not args.no_cpu}
                  if args.only in ("life512", "life3d", "energy", "batch",
                                   "refbench", "surrogate") else {})
        if args.leg_steps and args.only in ("life512", "life3d", "energy",
                                            "batch"):
            kwargs["steps"] = args.leg_steps
        if args.only == "energy":
            if args.fused_tile:
                kwargs["tiles"] = (tuple(int(v) for v in
                                         args.fused_tile.split("x")), )
            if args.fused_depth:
                kwargs["depths"] = (args.fused_depth, )
        if args.only == "batch" and args.bits_depth:
            kwargs["depth"] = args.bits_depth
            kwargs["strip_rows"] = args.bits_strips
        if args.only in ("energy", "batch") and args.no_plain:
            kwargs["plain_steps"] = 0
        print(json.dumps({args.only: _LEGS[args.only](**kwargs)}), flush=True)
        return 0

    from xentica_amd.runtime import hip

    device = hip.device()
    log("device: %s" % device.name)
    log("headline: GameOfLife 16384^2 through model.step(), warmup %d, "
        "steps %d" % (args.warmup, args.steps))
    model, ms = headline(args.steps, args.warmup)
    engine = model._engine
    n = model.cells_num
    value = n * args.steps / (ms / 1e3)
    log("%d steps in %.3f ms: %.4g cell-updates/s" % (args.steps, ms, value))
    crc = engine.crc32()
    total = args.warmup + args.steps
    plain_crc, plain_rate = plain_reference(
        total, 0 if args.no_plain else 20)
    verified = crc == plain_crc
    log("verification against the plain engine after %d steps: %s"
        % (total, verified))
    steady = steady_state(model) if args.steps < 2000 else None
    achieved = value * BYTES_PER_CELL_UPDATE / 1e9
    prof = counters()
    traffic = prof.get("hbm_bytes_per_launch")
    launch_s = ms / 1e3 / args.steps * engine.depth
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u1",
        "data": "synthetic",
        "config": {
            "workload": "life_b3s23_torus_16384x16384_bitpacked",
            "api": "GameOfLife(Life16kExperiment).step(), default engine "
                   "(bits)",
            "size": list(SIZE), "rule": "B3/S23", "seed_word": WORD,
            "density": 0.5, "temporal_depth": engine.depth,
            "row_strips": engine.strips or "library default (364 at "
                                           "16384^2: 91 workgroups of four "
                                           "strips handing rows over in "
                                           "LDS, eight 62-word tiles per "
                                           "strip, the 16-word remainder "
                                           "in segments shared three per "
                                           "wave, three waves per SIMD)",
            "layout": "64 cells per uint64 (32 per uint32 lane word)",
            "device": device.name,
        },
        "verified": verified,
        "crc_final": crc,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "hbm_frac": (traffic / launch_s / 1e9 / HBM_PEAK_GBS
                         if traffic else None),
            "valu_frac": (prof["sq_insts_valu_per_launch"] / launch_s
                          / VALU_ISSUE_PEAK
                          if prof.get("sq_insts_valu_per_launch") else None),
            "binding_resource": "VALU issue (temporal blocking: 8 steps per "
                                "HBM round trip)",
        },
    }
    if steady:
        result["steady_state"] = steady
    if plain_rate:
        result["plain_engine"] = plain_rate
        result["speedup_vs_plain"] = value / plain_rate["value"]
    if not args.no_512:
        log("Life 512^2 through model.step() ...")
        result["life512"] = life512(cpu=not args.no_cpu)
    if not args.no_render:
        log("render ...")
        result["render"] = render_leg()
    if not args.no_refbench:
        log("the reference's benchmark models ...")
        result["refbench"] = refbench(cpu=not args.no_cpu)
    if not args.no_3d:
        log("3-D workload ...")
        result["life3d"] = life3d(cpu=not args.no_cpu)
    if not args.no_energy:
        log("EnergyLife workload ...")
        result["energy_life"] = energy_life(cpu=not args.no_cpu)
    if not args.no_batch:
        log("batched workload ...")
        result["batched"] = batched(cpu=not args.no_cpu)
    if not args.no_surrogate:
        log("surrogate ...")
        result["surrogate"] = surrogate(cpu=not args.no_cpu)
    if not args.no_cpu:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)
    return 0 if verified else 1


if __name__ == "__main__":
    sys.exit(main())
