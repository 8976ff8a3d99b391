"""Launch-shape autotuner for the thermal Monte Carlo and the ray tracer.

The results of both do not depend on the launch shape (the thermal kernel is
bit-reproducible across block sizes and grid sizes, the images across ray
tiles), only the speed does, so the tuner simply times a few candidates on the
workload at hand and keeps the fastest:

* thermal: threads per block and resident blocks per CU of the persistent
  launch (``rt_thermal_opts.block_size`` / ``blocks_per_cu``), timed on a
  fraction of the run's packets;
* images: the ray tile one thread block traces (``rt_camera.tile_w`` /
  ``tile_h``), timed on forward + adjoint.

Choices are cached per workload signature for the life of the process.
"""
import time

THERMAL_CANDIDATES = (
    dict(block_size=256, blocks_per_cu=0),   # 0: as many as the occupancy allows
    dict(block_size=128, blocks_per_cu=0),
    dict(block_size=64, blocks_per_cu=0),
    dict(block_size=256, blocks_per_cu=2),
)
TILE_CANDIDATES = ((64, 1), (16, 4), (8, 8), (32, 8), (16, 16))

_cache = {}


def _best(key, candidates, timer):
    if key in _cache:
        return _cache[key]
    times = []
    for c in candidates:
        times.append((timer(c), c))
    best = min(times, key=lambda x: x[0])[1]
    _cache[key] = best
    _cache[key + ("times",)] = times
    return best


def timings(key):
    """(seconds, candidate) pairs measured for a signature, if tuned."""
    return _cache.get(tuple(key) + ("times",))


def tune_thermal(key, run, candidates=THERMAL_CANDIDATES):
    """run(**launch_opts) performs one (shortened) thermal run; returns the
    fastest launch options."""
    import torch

    def timer(c):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(**c)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    return _best(tuple(key), candidates, timer)


def tune_tiles(key, render, candidates=TILE_CANDIDATES, reps=3):
    """render(tile) traces one image (and its adjoint); returns the fastest tile."""
    import torch

    def timer(tile):
        render(tile)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            render(tile)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    return _best(tuple(key), candidates, timer)
