"""Autotuner of the batch engine: picks the round kernels' workgroup size, the
LDS traversal-stack depth per lane and the rounds per graph launch by timing
a calibration batch -- the job mix itself, scaled down -- on the device at
hand, and sizes the slot pool (the length of the ray queues) from an HBM
budget.

A coordinate search: kernel shape first (at the default rounds), then rounds
per launch.  Every candidate renders the same calibration jobs, so the times
compare like for like."""
from __future__ import annotations

import time
from dataclasses import dataclass, field, replace

from . import Batch, DeviceEnv, DeviceScene, RenderJob
from .workloads import JobSpec

SHAPES = [(b, s) for b in (64, 128, 256) for s in (8, 16, 24)]
ROUNDS = (8, 16, 32)


@dataclass
class BatchTuning:
    block_size: int = 256
    short_stack: int = 16
    rounds_per_launch: int = 16
    slot_capacity: int = 4 << 20
    trials: list = field(default_factory=list)  # (params, seconds)

    def kwargs(self) -> dict:
        return dict(block_size=self.block_size, short_stack=self.short_stack,
                    rounds_per_launch=self.rounds_per_launch, slot_capacity=self.slot_capacity)


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


def tune_batch(env: DeviceEnv, scenes: list[DeviceScene], jobs: list[JobSpec], max_trace_depth: int,
               log=None) -> BatchTuning:
    cal = calibration_jobs(jobs)
    best = BatchTuning()
    need = sum(((j.width + 7) // 8) * ((j.height + 7) // 8) * 64 for j in cal)
    best.slot_capacity = max(need, 1 << 16)

    def trial(t: BatchTuning) -> float:
        sec, rjs = run_batch(env, scenes, cal, max_trace_depth, **t.kwargs())
        for r in rjs:
            r.free()
        best.trials.append((t.kwargs(), round(sec, 4)))
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
    return best
