"""Host-side checks (no GPU) of the batched and serving workloads and of the
autotuner's sizing helpers: the generators produce the stated distributions,
deterministically per seed."""
import math

import numpy as np
import pytest

from crystalgraphics_amd import autotune, workloads


def test_batch_jobs_follow_the_stated_mix():
    jobs = workloads.batch_jobs(4000, 4, seed=3)
    w = np.array([j.width for j in jobs])
    h = np.array([j.height for j in jobs])
    assert w.min() >= 64 and w.max() <= 1920 and h.min() >= 64 and h.max() <= 1080
    # log-uniform: the exponent is uniform on [0, 1]
    t = np.log(w / 64) / np.log(1920 / 64)
    assert abs(t.mean() - 0.5) < 0.02 and abs(np.median(t) - 0.5) < 0.03
    assert set(j.spp for j in jobs) == {1, 4, 16, 64}
    d = np.array([j.depth for j in jobs])
    assert d.min() == 2 and d.max() == 12 and abs(d.mean() - 7) < 0.2
    assert set(j.scene for j in jobs) == {0, 1, 2, 3}
    for j in jobs[:200]:
        p, v = np.array(j.position), np.array(j.direction)
        assert abs(np.linalg.norm(p) - 3.0) < 1e-9 and np.allclose(v, -p / 3.0)
        assert abs(v[2]) < 0.9  # the camera basis uses dir x z
    again = workloads.batch_jobs(4000, 4, seed=3)
    assert [(j.width, j.spp, j.depth, j.position) for j in jobs] == [(j.width, j.spp, j.depth, j.position) for j in again]


def test_serve_requests_are_poisson_with_weighted_sizes():
    rate, dur = 50.0, 200.0
    reqs = workloads.serve_requests(rate, dur, 8, seed=5)
    t = np.array([r[0] for r in reqs])
    assert np.all(np.diff(t) > 0) and t[0] >= 0 and t[-1] < dur
    n = len(reqs)
    assert abs(n - rate * dur) < 5 * math.sqrt(rate * dur)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / rate) < 0.05 / rate
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05  # exponential gaps
    counts = [sum(1 for _, s in reqs if (s.width, s.height) == r) for r in workloads.SERVE_RESOLUTIONS]
    expect = np.array(workloads.SERVE_WEIGHTS) / sum(workloads.SERVE_WEIGHTS) * n
    chi2 = sum((c - e) ** 2 / e for c, e in zip(counts, expect))
    assert chi2 < 16.3  # 3 dof, p ~ 0.001
    assert all(s.spp in (1, 4, 16) and s.depth == 8 and 0 <= s.scene < 8 for _, s in reqs)


def test_calibration_jobs_cap_size_and_spp():
    jobs = workloads.batch_jobs(64, 4, seed=1)
    cal = autotune.calibration_jobs(jobs, max_pixels=640 * 360, max_spp=4)
    assert len(cal) == len(jobs)
    for a, b in zip(jobs, cal):
        assert b.width * b.height <= 640 * 360 and b.spp == min(a.spp, 4) and b.depth == a.depth
        if a.width * a.height <= 640 * 360:
            assert (b.width, b.height) == (a.width, a.height)
        else:
            assert abs(b.width / b.height - a.width / a.height) < 0.05


def test_slot_sizing():
    assert autotune.slot_bytes(12) == 64 + 32 + 4 + 8 + 128 * 11
    assert autotune.slot_bytes(1) == 64 + 32 + 4 + 8 + 128
    assert autotune.capacity_for(1 << 30, 8, 10 ** 9) == (1 << 30) // autotune.slot_bytes(8)
    assert autotune.capacity_for(1 << 30, 8, 5000) == 1 << 16  # the floor
    assert autotune.capacity_for(64 << 30, 12, 3_000_000) == 3_000_000


def test_pool_from_the_chosen_fraction():
    jobs = [workloads.JobSpec(0, 100, 50, 4, 8, (3, 0, 0), (-1, 0, 0)),
            workloads.JobSpec(1, 64, 64, 1, 2, (0, 3, 0), (0, -1, 0))]
    # 100x50 -> 13x7 tiles, 64x64 -> 8x8 tiles, 64 slots per tile
    assert autotune.resident_pixels(jobs) == (13 * 7 + 8 * 8) * 64
    t = autotune.BatchTuning(slot_fraction=0.5)
    assert t.capacity(jobs, 8, 64 << 30) == 1 << 16  # half of 9920 is under the floor
    big = [workloads.JobSpec(0, 1920, 1080, 1, 8, (3, 0, 0), (-1, 0, 0))] * 8
    need = autotune.resident_pixels(big)
    assert t.capacity(big, 8, 64 << 30) == need // 2
    assert t.capacity(big, 8, 1 << 30) == (1 << 30) // autotune.slot_bytes(8)  # the budget binds


def test_view_order():
    jobs = workloads.batch_jobs(16, 4, seed=3)
    order = autotune.view_order(jobs, "largest_first")
    assert sorted(order) == list(range(16))
    work = [jobs[i].pixels * jobs[i].spp * jobs[i].depth for i in order]
    assert work == sorted(work, reverse=True)
    assert autotune.view_order(jobs, "as_given") == list(range(16))
    with pytest.raises(ValueError):
        autotune.view_order(jobs, "random")


@pytest.mark.parametrize("q", [0, 1, 37.5, 50, 99, 100])
def test_percentile_matches_numpy(q):
    v = list(np.random.default_rng(2).random(101))
    assert abs(workloads.percentile(v, q) - np.percentile(v, q)) < 1e-12


def test_scene_triangle_counts():
    assert workloads.scene_triangles(1.0) == 16 * 131072 + 4096
    tri = {k: workloads.scene_triangles(v) for k, v in workloads.SCENE_SCALES.items()}
    assert abs(tri["0.5M"] - 0.5e6) < 0.03e6 and abs(tri["4M"] - 4e6) < 0.1e6
