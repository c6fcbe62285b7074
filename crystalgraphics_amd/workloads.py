"""Seeded synthetic workloads of the batched and serving forms: render-job
specifications over several resident scenes built with the 16-sphere
generator ("c1") at different triangle counts.  Pure host-side data; the
rendering goes through Batch / RenderJob."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .native import Camera, Conf

# Sphere subdivision factor of the "c1" generator (n = 128 * f, 8 n^2
# triangles per sphere, 16 spheres) for ~0.5M, 1M, 2M and 4M triangles.
SCENE_SCALES = {"0.5M": 62 / 128, "1M": 88 / 128, "2M": 1.0, "4M": 176 / 128}


def scene_triangles(scale: float) -> int:
    n = int(128 * scale)
    return 16 * 8 * n * n + 4096


@dataclass
class JobSpec:
    scene: int          # index into the batch's scene list
    width: int
    height: int
    spp: int
    depth: int
    position: tuple
    direction: tuple

    def conf(self) -> Conf:
        return Conf.make(self.width, self.height, self.spp, self.depth)

    def camera(self) -> Camera:
        return Camera.make(self.position, self.direction, (0.5 * self.width / self.height, 0.5))

    @property
    def pixels(self) -> int:
        return self.width * self.height


def orbit_camera(rng: np.random.Generator, radius: float) -> tuple[tuple, tuple]:
    """A camera on a sphere of `radius` aimed at the origin, away from the
    poles (the camera basis uses dir x z, which vanishes there)."""
    while True:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if abs(v[2]) < 0.9:
            break
    pos = radius * v
    return tuple(float(x) for x in pos), tuple(float(x) for x in -v)


def batch_jobs(n_jobs: int = 64, n_scenes: int = 4, seed: int = 1, min_res=(64, 64), max_res=(1920, 1080),
               spp_choices=(1, 4, 16, 64), depth_range=(2, 12), radius: float = 3.0) -> list[JobSpec]:
    """The batched form's job mix: resolution log-uniform between min_res and
    max_res (one uniform exponent for both axes), spp uniform over
    spp_choices, depth uniform on depth_range, camera on a sphere of
    `radius` aimed at the origin, scene uniform over the resident scenes."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n_jobs):
        t = rng.uniform()
        w = int(round(min_res[0] * (max_res[0] / min_res[0]) ** t))
        h = int(round(min_res[1] * (max_res[1] / min_res[1]) ** t))
        pos, d = orbit_camera(rng, radius)
        jobs.append(JobSpec(int(rng.integers(n_scenes)), w, h, int(rng.choice(spp_choices)),
                            int(rng.integers(depth_range[0], depth_range[1] + 1)), pos, d))
    return jobs


SERVE_RESOLUTIONS = ((256, 256), (640, 360), (1280, 720), (1920, 1080))
SERVE_WEIGHTS = (4, 3, 2, 1)


def serve_requests(rate: float, duration: float, n_scenes: int, seed: int = 1, spp_choices=(1, 4, 16),
                   depth: int = 8, radius: float = 3.0) -> list[tuple[float, JobSpec]]:
    """Poisson arrivals at `rate` per second over `duration` seconds; each
    request: a scene id, a camera on the orbit sphere, a resolution from
    SERVE_RESOLUTIONS weighted SERVE_WEIGHTS, spp from spp_choices."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    p = np.array(SERVE_WEIGHTS, np.float64)
    p /= p.sum()
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        w, h = SERVE_RESOLUTIONS[int(rng.choice(len(SERVE_RESOLUTIONS), p=p))]
        pos, d = orbit_camera(rng, radius)
        out.append((t, JobSpec(int(rng.integers(n_scenes)), w, h, int(rng.choice(spp_choices)), depth, pos, d)))
    return out


def percentile(values, q: float) -> float:
    v = sorted(values)
    if not v:
        return math.nan
    k = (len(v) - 1) * q / 100.0
    lo, hi = math.floor(k), math.ceil(k)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)
