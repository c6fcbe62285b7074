"""Autotuner of the batch engines.  For the wavefront engine it picks the
round kernels' workgroup size, the LDS traversal-stack depth per lane, the
rounds per graph launch and the size of the slot pool (the length of the ray
queues, as a fraction of the pixels the jobs could keep resident); for the
one-launch per-pixel batch it picks the order of the jobs' tiles in the work
queue.  Every candidate renders the same calibration batch -- the job mix
itself, scaled down -- on the device at hand, so the times compare like for
like.

A coordinate search: kernel shape first (at the default rounds and a pool
that holds every calibration pixel), then rounds per launch, then the pool
fraction.  The run applies the chosen fraction to its own jobs
(BatchTuning.capacity), within an HBM budget."""
from __future__ import annotations

import time
from dataclasses import dataclass, field, replace

from . import Batch, DeviceEnv, DeviceScene, RenderJob, ViewBatch
from .workloads import JobSpec

SHAPES = [(b, s) for b in (64, 128, 256) for s in (8, 16, 24)]
ROUNDS = (8, 16, 32)
SLOT_FRACTIONS = (0.125, 0.25, 0.5, 1.0)  # of the pixels the jobs could keep resident
ORDERS = ("largest_first", "as_given")     # tile order of the one-launch per-pixel batch


@dataclass
class BatchTuning:
    block_size: int = 256
    short_stack: int = 16
    rounds_per_launch: int = 16
    slot_capacity: int = 4 << 20
    slot_fraction: float = 1.0          # pool size / pixels the jobs could keep resident
    view_order: str = "largest_first"   # one-launch per-pixel batch: order of the jobs' tiles
    trials: list = field(default_factory=list)  # (params, seconds)

    def kwargs(self) -> dict:
        return dict(block_size=self.block_size, short_stack=self.short_stack,
                    rounds_per_launch=self.rounds_per_launch, slot_capacity=self.slot_capacity)

    def capacity(self, jobs: list[JobSpec], max_trace_depth: int, budget_bytes: int) -> int:
        """The pool for `jobs` at the chosen fraction, within the budget."""
        return capacity_for(budget_bytes, max_trace_depth, int(self.slot_fraction * resident_pixels(jobs)))


def resident_pixels(jobs: list[JobSpec]) -> int:
    """Slots that would hold every pixel of `jobs` at once (tiles of 8x8)."""
    return sum(((j.width + 7) // 8) * ((j.height + 7) // 8) * 64 for j in jobs)


def view_order(jobs: list[JobSpec], order: str) -> list[int]:
    """Job indices in the tile order of the one-launch per-pixel batch:
    largest_first puts the jobs with the most queries (pixels x spp x depth)
    first, so the longest finish last the least; as_given keeps submission
    order."""
    if order == "largest_first":
        return sorted(range(len(jobs)), key=lambda i: -jobs[i].pixels * jobs[i].spp * jobs[i].depth)
    if order == "as_given":
        return list(range(len(jobs)))
    raise ValueError(f"unknown order {order!r}")


def calibration_jobs(jobs: list[JobSpec], max_pixels: int = 640 * 360, max_spp: int = 4) -> list[JobSpec]:
    """The job mix scaled down: resolution capped at max_pixels (aspect kept),
    spp capped at max_spp."""
    out = []
    for j in jobs:
        w, h = j.width, j.height
        if w * h > max_pixels:
            f = (max_pixels / (w * h)) ** 0.5
            w, h = max(8, int(w * f)), max(8, int(h * f))
        out.append(replace(j, width=w, height=h, spp=min(j.spp, max_spp)))
    return out


def run_batch(env: DeviceEnv, scenes: list[DeviceScene], jobs: list[JobSpec], max_trace_depth: int,
              **kwargs) -> tuple[float, list[RenderJob]]:
    """Render `jobs` through one Batch, admitting as slots free up; returns
    the wall time and the finished RenderJobs (buffers still allocated)."""
    b = Batch(env, max_trace_depth=max_trace_depth, **kwargs)
    rjs = [RenderJob(env, scenes[j.scene], j.camera(), j.conf()) for j in jobs]
    try:
        env.sync()
        t0 = time.perf_counter()
        pending = list(range(len(rjs)))
        finished = 0
        while pending or finished < len(rjs):
            pending = [i for i in pending if not b.submit(rjs[i])]
            b.launch(1)
            done, _ = b.poll()
            finished += len(done)
        return time.perf_counter() - t0, rjs
    finally:
        b.close()


def slot_bytes(max_trace_depth: int) -> int:
    """HBM per pixel slot: path state 64 B, hit 32 B, job id 4 B, two queue
    entries 8 B, and one 128-B transport frame per level below the deepest."""
    return 64 + 32 + 4 + 8 + 128 * max(1, max_trace_depth - 1)


def capacity_for(budget_bytes: int, max_trace_depth: int, need: int) -> int:
    """Slots for an HBM budget, no more than the jobs can use."""
    return max(1 << 16, min(int(budget_bytes // slot_bytes(max_trace_depth)), need))


def run_view_batch(env: DeviceEnv, scenes: list[DeviceScene], jobs: list[JobSpec], order: str) -> float:
    """Render `jobs` in one launch of the per-pixel kernel in the given tile
    order; returns the wall time."""
    rjs = [RenderJob(env, scenes[j.scene], j.camera(), j.conf()) for j in jobs]
    try:
        env.sync()
        t0 = time.perf_counter()
        vb = ViewBatch(env, [rjs[i] for i in view_order(jobs, order)])
        vb.wait()
        sec = time.perf_counter() - t0
        vb.release()
        return sec
    finally:
        for r in rjs:
            r.free()


def tune_batch(env: DeviceEnv, scenes: list[DeviceScene], jobs: list[JobSpec], max_trace_depth: int,
               log=None) -> BatchTuning:
    cal = calibration_jobs(jobs)
    best = BatchTuning()
    need = resident_pixels(cal)
    best.slot_capacity = max(need, 1 << 16)

    def trial(t: BatchTuning) -> float:
        sec, rjs = run_batch(env, scenes, cal, max_trace_depth, **t.kwargs())
        for r in rjs:
            r.free()
        best.trials.append((dict(t.kwargs(), slot_fraction=t.slot_fraction), round(sec, 4)))
        if log:
            log(f"autotune {t.kwargs()}: {sec * 1e3:.1f} ms")
        return sec

    trial(best)  # warm-up (first-use costs)
    best_t = trial(best)
    for b, s in SHAPES:
        if (b, s) == (best.block_size, best.short_stack):
            continue
        t = trial(replace(best, block_size=b, short_stack=s, trials=[]))
        if t < best_t:
            best_t, best.block_size, best.short_stack = t, b, s
    for r in ROUNDS:
        if r == best.rounds_per_launch:
            continue
        t = trial(replace(best, rounds_per_launch=r, trials=[]))
        if t < best_t:
            best_t, best.rounds_per_launch = t, r
    for f in SLOT_FRACTIONS:
        if f == best.slot_fraction:
            continue
        t = trial(replace(best, slot_fraction=f, slot_capacity=max(int(f * need), 1 << 16), trials=[]))
        if t < best_t:
            best_t, best.slot_fraction, best.slot_capacity = t, f, max(int(f * need), 1 << 16)
    order_t = {}
    run_view_batch(env, scenes, cal, ORDERS[0])  # warm-up (first-use costs)
    for o in ORDERS:
        order_t[o] = run_view_batch(env, scenes, cal, o)
        best.trials.append(({"view_order": o}, round(order_t[o], 4)))
        if log:
            log(f"autotune view_order={o}: {order_t[o] * 1e3:.1f} ms")
    best.view_order = min(order_t, key=order_t.get)
    return best
