"""Benchmark: thermal-MC photon packets/s (+ image rays/s with adjoint) on one MI355X.

Workload (the production-size disk + envelope): spherical grid 256 x 128 x 1
cells (0.1-5000 au, equatorial mirror), flared disk (1e-3 Msun, Sigma ~ r^-0.5)
plus an Ulrich envelope (0.1 Msun, Rc = 50 au, 20 deg cavity), two dust species
with kappa_abs = 1e3 (lam/um)^-1 (albedo 0.5, isotropic), a 4000 K 2 Rsun star,
100 wavelengths, modified random walk on (pdspy's fitting setting), nphot = 1e8
packets per run, seeds 0-3.  One step = one thermal run (seed = step mod 4, so
consecutive steps cover the four seeds) plus the copy of its per-species
temperatures to the host (what run_thermal hands back).  Steps run on
--streams (default 4) host threads, each with its own context, HIP stream and
copy of the model, step k on thread k mod 4: the runs share the GPU, so one
run's last long random walks (its tail, when few of its waves still hold
packets) overlap the others' bulk (each run's temperatures are bit-identical
to a run on its own).  A warm-up step is a run of another seed.  No timed step is started once it would end past
--step-budget-s seconds after the process started; `steps` reports the steps
actually timed.

After the timed steps (each leg only while time is left): the 256 x 256 1.3 mm
image at 60 deg (2" field at 140 pc) of the resulting temperatures and its
adjoint, without and with the scattering source, and its visibilities; the
GPU against the FP64 oracle on the same workload at equal nphot and seed
(`accuracy`); the CPU baseline (the oracle on the host's cores, a bounded
sample); the workload's four seeds in one batched launch, timed once
(`batched_seeds`).

Prints ONE JSON line.  --gpus above 1 is not supported (one GPU per box).

--batch: instead, the batched sampler step -- 64 disk+envelope models drawn as
configs.sampler_batch (uneven grids 64-256 x 64-128, 1e6-1e7 packets each, MRW
on), each step = one thermal launch for all of them, then per model the 1.3 mm
256x256 image, visibilities at 5e4 (u,v) points, chi^2 and the adjoint
gradient, replayed from a hipGraph -- timed against the same work done model by
model through the single-model calls.

--vis: instead, the two visibility samplers (exact DFT; galario's method, the
image's transform on its pixel grid + bilinear interpolation) on 256^2 and
1024^2 images at 5e4 and 1e6 (u,v) points.
"""
import argparse
import hashlib
import json
import os
import sys
import time

T_PROC0 = time.perf_counter()

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_VECTOR_TFLOPS = 78.6  # 256 CU x 4 SIMD x 16 lanes x 2 flop x 2.4 GHz, no packed FP32
STAGGER_S = 0.1  # least time between two thermal launches of the overlapped steps
CALIB_BYTES = 256 << 20  # device copy whose FETCH_SIZE / WRITE_SIZE scale the counters


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--nphot", type=float, default=1e8)
    p.add_argument("--seeds", type=int, default=4,
                   help="runs of the side measurement that times the workload's seeds 0..N-1 "
                        "in one batched launch; 1 = no such measurement")
    p.add_argument("--step-budget-s", type=float, default=470.0,
                   help="no timed step is started once it would end later than this many "
                        "seconds after the process started")
    p.add_argument("--side-budget-s", type=float, default=540.0,
                   help="the batched four-seed launch runs only if it ends before this")
    p.add_argument("--stagger-s", type=float, default=STAGGER_S,
                   help="least time between two thermal launches of the overlapped steps")
    p.add_argument("--streams", type=int, default=4,
                   help="steps run on this many host threads / HIP streams, overlapping "
                        "one run's last random walks with the next run (1 = back to back)")
    p.add_argument("--acc-nphot", type=float, default=2e5,
                   help="packets of the GPU-vs-oracle accuracy comparison")
    p.add_argument("--max-events", type=float, default=1e6)
    p.add_argument("--mrw", type=int, default=1)
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="target wall time of the bounded CPU sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-image", action="store_true")
    p.add_argument("--nphot-scat", type=float, default=1e5,
                   help="scattering-source packets per image (pdspy's fits: 1e5)")
    p.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "thermal_traffic.json"))
    p.add_argument("--batch", action="store_true", help="batched sampler step instead")
    p.add_argument("--batch-models", type=int, default=64)
    p.add_argument("--batch-k", type=int, default=50000)
    p.add_argument("--no-sequential", action="store_true")
    p.add_argument("--dust", action="store_true",
                   help="batch: each model's disk dust drawn from a DustGenerator (a_max, p), "
                        "drawn again with the new parameters")
    p.add_argument("--vis", action="store_true",
                   help="visibility samplers only: exact DFT and galario's method")
    return p.parse_args()


def vis_main(a):
    """Visibilities of a resolved source (Gaussian of sigma 0.15 arcsec, offset
    from the centre) at 256^2 and 1024^2 pixels, at the workload's 5e4 (u,v)
    points and at 1e6: the exact DFT (rt_vis_dft) and galario's method
    (rt_vis_grid, grid transform + bilinear), with the latter's deviation from
    the former."""
    import torch
    from pdspy_amd import configs, inputs, rt
    rows = []
    for n in (256, 1024):
        pix = 2.0 / n  # 2" field
        x = (np.arange(n) - n // 2) * pix
        X, Y = np.meshgrid(x - 0.05, x + 0.02)
        img_h = np.exp(-(X ** 2 + Y ** 2) / (2 * 0.15 ** 2))
        img = torch.tensor(img_h, dtype=torch.float32, device="cuda")
        dxy = pix * inputs.arcsec
        for K in (50000, 1000000):
            u, v = configs.uv_points(K=K, lam_m=1.3e-3, seed=0)
            u = torch.tensor(u, device="cuda")
            v = torch.tensor(v, device="cuda")
            row = {"npix": n, "K": K}
            res = {}
            for name, fn in (("dft", rt.vis_dft), ("grid", rt.vis_grid)):
                fn(img, dxy, u, v)
                torch.cuda.synchronize()
                reps = 3 if (name == "dft" and n * n * K > 1e11) else 10
                t0 = time.perf_counter()
                for _ in range(reps):
                    res[name] = fn(img, dxy, u, v)
                torch.cuda.synchronize()
                row[name + "_ms"] = (time.perf_counter() - t0) / reps * 1e3
            d = torch.hypot(res["grid"][0] - res["dft"][0], res["grid"][1] - res["dft"][1])
            row["grid_max_dev_over_V0"] = float(d.max()) / float(img_h.sum())
            row["speedup"] = row["dft_ms"] / row["grid_ms"]
            print("vis: %s" % row, file=sys.stderr, flush=True)
            rows.append(row)
    out = {"metric": "visibilities/s (galario's method, 256x256 image, 5e4 (u,v))",
           "value": 50000 / (rows[0]["grid_ms"] * 1e-3), "unit": "visibilities/s", "n_gpus": 1,
           "steps": 10, "warmup": 1, "ms_per_step": rows[0]["grid_ms"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": "Gaussian disk image, 2 arcsec field, (u,v) of 15-2000 m "
                                  "baselines at 1.3 mm"},
           "cases": rows}
    print(json.dumps(out), flush=True)


def batch_main(a):
    import torch
    from pdspy_amd import configs
    from pdspy_amd.sampler import SamplerStep

    params = configs.sampler_batch(a.batch_models, seed=0, dust=a.dust)
    gen = configs.synthetic_dust_generator() if a.dust else None
    t0 = time.perf_counter()
    s = SamplerStep(params, npix=256, K=a.batch_k, max_events=int(a.max_events), mrw=a.mrw,
                    dust_gen=gen)
    t_build = time.perf_counter() - t0
    tuned = s.autotune()
    print("batch: built and tuned", json.dumps(tuned), file=sys.stderr, flush=True)
    for w in range(max(1, a.warmup)):
        s.step(seed=100 + w, use_graph=True)       # first one captures the graph
    torch.cuda.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for k in range(a.steps):
        stats, ms = s.step(seed=k, use_graph=True)
        dev_ms.append(ms)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / a.steps
    print("batch: %d steps, %.2f s each" % (a.steps, wall), file=sys.stderr, flush=True)
    # a new parameter draw for the same grids: densities rebuilt on the device,
    # then a graph step with them; the host builders timed on 8 of the models
    # (the first draw also makes the models' update buffers and per-temperature
    # tables; the second is what every later draw costs)
    draws = []
    for seed in (1, 2):
        new = configs.sampler_batch(a.batch_models, seed=seed, dust=a.dust)
        for p, q in zip(new, params):
            p.update(nr=q["nr"], nt=q["nt"], nphot=q["nphot"])
        draws.append(s.set_params(new))
    t_first, t_gen = draws
    t0 = time.perf_counter()
    stats_new, _ = s.step(seed=7, use_graph=True)
    torch.cuda.synchronize()
    t_new = time.perf_counter() - t0
    chi_new = [float(c.item()) for c in s.chi2]
    t0 = time.perf_counter()
    for p in new[:8]:
        configs.batch_model(p)
    t_host_gen = (time.perf_counter() - t0) / 8
    print("batch: new parameters: device update %.1f ms for %d models (first draw %.1f ms), "
          "step %.2f s; host builders %.2f s per model" % (t_gen * 1e3, s.n, t_first * 1e3,
                                                          t_new, t_host_gen),
          file=sys.stderr, flush=True)
    seq = None
    if not a.no_sequential:
        seq, _ = s.sequential(seed=0, progress=lambda i, t: print(
            "batch: sequential model %d done at %.1f s" % (i, t), file=sys.stderr, flush=True))
        print("batch: sequential %.2f s" % seq, file=sys.stderr, flush=True)
    npk = int(s.nphot.sum())
    out = {
        "metric": "batched sampler step (thermal MC + 1.3 mm image + visibilities + chi^2 "
                  "gradient), %d models%s" % (s.n, ", per-model dust" if a.dust else ""),
        "value": s.n / wall, "unit": "models/s", "n_gpus": 1, "steps": a.steps,
        "warmup": max(1, a.warmup), "ms_per_step": wall * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "configs.sampler_batch(%d, seed=0): grids nr in {64,128,256} x "
                               "ntheta in {64,128}, nphot 1e6-1e7 log-uniform, MRW %s; 256x256 "
                               "images at 1.3 mm, K=%d" % (s.n, "on" if a.mrw else "off", s.K),
                   "models": s.n, "packets_per_step": npk, "cells": [m.ncell for m in s.models]},
        "step_device_ms": dev_ms, "graph": True, "tuned": tuned,
        "packets_per_s": npk / wall, "build_s": t_build,
        "sequential_s": seq, "speedup_vs_sequential": (seq / wall) if seq else None,
        "killed": int(sum(x["n_killed"] for x in stats)),
        "new_params": {"device_density_ms": t_gen * 1e3, "first_draw_ms": t_first * 1e3,
                       "step_s": t_new,
                       "includes_dust_tables": bool(a.dust),
                       "host_builder_s_per_model": t_host_gen,
                       "chi2_finite": bool(np.all(np.isfinite(chi_new))),
                       "killed": int(sum(x["n_killed"] for x in stats_new))},
    }
    print(json.dumps(out), flush=True)


def dT_stats(T_ref, T, visits, min_visits=100):
    """median and 99th percentile of |dT|/T over cells with >= min_visits entries
    and a positive reference temperature (all species)."""
    sel = (np.asarray(visits) >= min_visits)[None, :] & (T_ref > 0)
    d = np.abs(T[sel] - T_ref[sel]) / T_ref[sel]
    return float(np.median(d)), float(np.percentile(d, 99)), int(sel.sum())


def elapsed():
    """seconds since this process started (the driver's clock starts there)"""
    return time.perf_counter() - T_PROC0


def host_cpu():
    """The host's CPU as the baseline sees it: logical CPUs, the CPUs this
    process may run on, the threads the baseline uses (OMP_NUM_THREADS when
    the environment sets it, else every allowed CPU) and the model name."""
    nproc = os.cpu_count()
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = allowed
    if omp and omp.isdigit() and int(omp) > 0:
        threads = min(int(omp), allowed)
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": nproc, "allowed_cpus": allowed, "threads": threads,
            "omp_num_threads_env": omp, "model_name": model}


def cpu_baseline(arr, max_events, target_s, mrw, cpu):
    """The FP64 oracle (same physics) on the host's cores: immediate
    re-emission, packets free-running over OpenMP threads that share the
    tallies (no epoch barriers), on a bounded sample of the same workload.
    This is a CPU restatement of the algorithm, not RADMC-3D (which is not
    installed on the GPU host)."""
    import oracle
    nthr = cpu["threads"]
    n = 10000
    t0 = time.perf_counter()
    oracle.thermal(arr, n, seed=100, max_events=int(max_events), lagged=False, nthreads=nthr,
                   mrw=bool(mrw))
    dt = time.perf_counter() - t0
    n2 = int(max(n, n * target_s / max(dt, 1e-3)))
    t0 = time.perf_counter()
    _, _, st = oracle.thermal(arr, n2, seed=101, max_events=int(max_events), lagged=False,
                              nthreads=nthr, mrw=bool(mrw))
    dt = time.perf_counter() - t0
    steps = st["n_cross"] + st["n_events"]
    return {"value": n2 / dt, "unit": "packets/s", "cores": nthr, "kind": "port",
            "sample": "CPU restatement, not RADMC-3D: %d packets (seed 101) of the same "
                      "workload through the FP64 oracle, immediate re-emission, OpenMP over "
                      "packets on %d threads, %.1f s wall, %.0f steps (crossings + events) "
                      "per packet" % (n2, nthr, dt, steps / n2),
            "steps_per_s": steps / dt, "steps_per_packet": steps / n2,
            "host": cpu}


def accuracy_vs_oracle(dm, arr, nphot, seed, max_events, mrw, nthreads):
    """GPU run and FP64 oracle run of the same workload at equal nphot and
    seed (same Philox streams and epoch schedule): median and 99th percentile
    of |dT|/T over cells with >= 100 visits."""
    import oracle
    dm.thermal(nphot, seed, count_visits=1, max_events=int(max_events), mrw=mrw)
    Tg, vis = dm.temperature(with_visits=True)
    t0 = time.perf_counter()
    To, vo, so = oracle.thermal(arr, nphot, seed=seed, max_events=int(max_events), lagged=True,
                                nthreads=nthreads, mrw=bool(mrw))
    t_or = time.perf_counter() - t0
    med, p99, ncell = dT_stats(To, Tg, vo)
    return {"median_dT_over_T": med, "p99_dT_over_T": p99, "cells_ge_100_visits": ncell,
            "nphot": int(nphot), "seed": int(seed), "oracle_s": t_or,
            "compared": "GPU vs FP64 oracle (lagged re-emission, the GPU's algorithm) on the "
                        "same workload at equal nphot and seed"}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    if a.batch:
        batch_main(a)
        return
    if a.vis:
        vis_main(a)
        return
    import torch
    from pdspy_amd import configs, inputs, rt

    def log(msg):
        print("bench [%6.1f s]: %s" % (elapsed(), msg), file=sys.stderr, flush=True)

    nphot = int(a.nphot)
    model = configs.disk_envelope()
    arr = rt.model_arrays(model.grid)
    dm = rt.DeviceModel(arr)
    S = dm.nspec
    opts = dict(max_events=int(a.max_events), mrw=a.mrw)
    log("model built (%d cells, %d species)" % (dm.ncell, S))

    # calibration copy of known size (scales FETCH_SIZE / WRITE_SIZE counters
    # when this program runs under rocprofv3 --pmc)
    src = torch.ones(CALIB_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst

    # steps run on `streams` host threads, each with its own context, stream and
    # copy of the model: step k goes to lane k mod streams.  The launches share
    # the GPU (with several lanes each leaves a CU per XCD free, where the other
    # lanes' short kernels run at once), so a run's tail of long random walks
    # overlaps the other runs' bulk.
    from pdspy_amd import _lib
    nl = max(1, int(a.streams))
    lane_opts = dict(opts, free_cus_per_xcd=1 if nl > 1 else 0)
    lanes = [(dm, torch.cuda.current_stream())]
    for _ in range(nl - 1):
        cx = _lib.Context(0)
        lanes.append((rt.DeviceModel(arr, ctx=cx), torch.cuda.Stream()))

    import threading
    launch_mu = threading.Lock()
    last_launch = [-1.0]

    def stagger():
        """A run is launched at least --stagger-s after the previous one (the
        lanes' launches do not all start in the same instant)."""
        with launch_mu:
            wait = last_launch[0] + a.stagger_s - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            last_launch[0] = time.perf_counter()

    def run_steps(nsteps, seed_of, budget, tag, first_visits=False):
        """nsteps runs over the lanes; returns per step (stats, T, seconds),
        None for a step not started because of the budget"""
        out = [None] * nsteps
        err = []

        def worker(li):
            dml, stream = lanes[li]
            last = None
            try:
                with torch.cuda.stream(stream):
                    for k in range(li, nsteps, nl):
                        if budget and last is not None and elapsed() + 1.2 * last > budget:
                            log("%s: lane %d stops before step %d (budget)" % (tag, li, k + 1))
                            return
                        stagger()
                        ts = time.perf_counter()
                        cv = 1 if (first_visits and k == 0) else 0
                        st = dml.thermal(nphot, seed_of(k), count_visits=cv, **lane_opts)
                        res = dml.temperature(with_visits=bool(cv))
                        last = time.perf_counter() - ts
                        out[k] = (st, res, last)
                        log("%s %d/%d (lane %d): seed %d, %.0f ms"
                            % (tag, k + 1, nsteps, li, seed_of(k), last * 1e3))
            except Exception as e:  # re-raised in the main thread
                err.append(e)

        th = [threading.Thread(target=worker, args=(li,)) for li in range(nl)]
        for t in th:
            t.start()
            time.sleep(0.01)  # lane 0 launches first
        for t in th:
            t.join()
        if err:
            raise err[0]
        return out

    # warm-up: runs of other seeds (the first also counts cell visits for the
    # accuracy mask)
    vis = None
    wres = run_steps(a.warmup, lambda k: 1000 + k, None, "warm-up", first_visits=True)
    if a.warmup > 0:
        vis = wres[0][1][1]
    torch.cuda.synchronize()

    # timed steps: step k = one run of seed k mod 4 (the workload's seeds 0-3)
    # plus the copy of its per-species temperatures to the host; no step is
    # started once it would end past the step budget
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = run_steps(a.steps, lambda k: k % 4, a.step_budget_s, "step")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    done = [r for r in res if r is not None]
    stats = [r[0] for r in done]
    temps = [r[1] for r in done]
    steps = len(done)
    seeds_done = [k % 4 for k, r in enumerate(res) if r is not None]
    value = nphot * steps / wall
    last = max(r[2] for r in done)

    # roofline of the MC kernel: algorithmic bytes B_alg = N_cross * 4 S +
    # N_abs * 16 per launch over the launch's average duration (HIP events on
    # the stream it runs on, the figure rocprofv3's kernel trace reports).
    # With several lanes the launches overlap and a duration includes the time
    # a launch shares the GPU with its neighbour; the timed launches'
    # algorithmic bytes over the timed wall time are reported beside it
    kern_ms = sum(st["kernel_ms"] for st in stats) / steps
    b_alg = sum(x["n_cross"] * 4 * S + x["n_abs"] * 16 for x in stats) / steps
    achieved = b_alg / (kern_ms * 1e-3) / 1e9
    achieved_wall = b_alg * steps / wall / 1e9
    traffic, traffic_src = None, None
    if os.path.exists(a.traffic_json):
        try:
            tj = json.load(open(a.traffic_json))
            if int(tj.get("nphot", -1)) == nphot and int(tj.get("mrw", -1)) == a.mrw:
                traffic = tj.get("bytes_per_launch")
                traffic_src = ("not measured in this run: rocprofv3 --pmc FETCH_SIZE / "
                               "WRITE_SIZE of the same kernel and workload, scaled by a "
                               "256 MiB copy, from %s" % os.path.relpath(a.traffic_json, ROOT))
        except (OSError, ValueError):
            traffic = None
    lane_steps = sum(x["n_cross"] + x["n_events"] for x in stats)
    wave_iters = sum(x["wave_iters"] for x in stats)
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
            "kernel": "k_thermal", "b_alg_per_launch": b_alg, "kernel_ms_per_launch": kern_ms,
            "launches_per_step": 1, "lanes": nl,
            "timing": ("HIP events on the launch's stream" if nl == 1 else
                       "HIP events on each launch's stream; the launches overlap on %d "
                       "streams, so a duration includes the time shared with the "
                       "neighbouring launch (achieved_over_wall: the timed launches' "
                       "algorithmic bytes over the timed wall time)" % nl),
            "achieved_over_wall": achieved_wall,
            "steps_per_s": lane_steps / (kern_ms * steps * 1e-3),
            "lane_utilisation": lane_steps / (64.0 * max(wave_iters, 1))}

    med = p99 = ncell = None
    if vis is not None and len(temps) >= 2:
        med, p99, ncell = dT_stats(temps[0], temps[1], vis)
    h0 = hashlib.sha256(temps[0].astype(np.float32).tobytes()).hexdigest()[:16]
    repeat_equal = None
    if len(temps) > 4:
        repeat_equal = bool(np.array_equal(temps[0], temps[4]))

    out = {
        "metric": "photon packets/s (thermal MC) + image rays/s w/ adjoint, at <=1% median dT/T",
        "value": value, "unit": "packets/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "disk+envelope 256x128x1 cells, 2 species, 100 lambda, "
                               "nphot=%g per run, one run per step (seed = step mod 4), "
                               "MRW %s, steps on %d streams" % (nphot, "on" if a.mrw else "off",
                                                                nl),
                   "nphot": nphot, "seeds": sorted(set(seeds_done)), "streams": nl,
                   "cells": dm.ncell, "species": S, "max_events": int(a.max_events),
                   "mrw": a.mrw, "steps_requested": a.steps},
        "roofline": roof,
        "seed_noise": {"median_dT_over_T": med, "p99_dT_over_T": p99,
                       "cells_ge_100_visits": ncell,
                       "compared": "timed runs seed 0 vs seed 1 at equal nphot: Monte Carlo "
                                   "noise between seeds, not an accuracy against a reference",
                       "T_hash_seed0": h0, "seed0_repeat_bitwise_equal": repeat_equal},
        "mc_stats": {k: stats[-1][k] for k in ("n_cross", "n_abs", "n_scat", "n_killed",
                                               "n_escape", "n_events", "n_mrw", "n_epochs",
                                               "n_nudged", "max_packet_steps")},
    }
    log("timed %d steps: %.3e packets/s" % (steps, value))
    cpu = host_cpu()

    if not a.no_image:
        # 1.3 mm, 256x256, 2" field at 140 pc, 60 deg; forward and adjoint
        T = torch.tensor(rt.imaging_temperature(_with_T(model, temps[0], dm)),
                         dtype=torch.float32, device="cuda")
        from pdspy_amd import tune
        pix_cm = 2.0 / 256 * 140 * inputs.AU
        dimg = torch.ones((256, 256), dtype=torch.float32, device="cuda")

        def render(tile):
            c = dm.camera(256, pix_cm, 1300.0, incl=60.0, tile=tile)
            dm.image(c, T)
            dm.image_adjoint(c, T, None, dimg)

        tile = tune.tune_tiles(("bench", 256, dm.nr, dm.nt, dm.np), render)
        cam = dm.camera(256, pix_cm, 1300.0, incl=60.0, tile=tile)
        for _ in range(2):
            dm.image(cam, T)
            dm.image_adjoint(cam, T, None, dimg)
        torch.cuda.synchronize()
        n_rep = 10
        t0 = time.perf_counter()
        for _ in range(n_rep):
            dm.image(cam, T)
        torch.cuda.synchronize()
        t_f = (time.perf_counter() - t0) / n_rep
        t0 = time.perf_counter()
        for _ in range(n_rep):
            dm.image(cam, T)
            dm.image_adjoint(cam, T, None, dimg)
        torch.cuda.synchronize()
        t_fa = (time.perf_counter() - t0) / n_rep
        # cells crossed by one image's rays and by one adjoint's reverse sweeps
        dm.image_crossings(reset=True)
        dm.image(cam, T)
        n_img, _ = dm.image_crossings(reset=True)
        dm.image_adjoint(cam, T, None, dimg)
        _, n_adj = dm.image_crossings(reset=True)
        # algorithmic bytes: k_image reads alpha and S (8 B) per crossed cell
        # and writes 4 B per pixel; the adjoint reads them twice (forward and
        # reverse sweep), reads the cotangent image and writes two FP64
        # gradients per cell
        npx, ncell_img = 65536, dm.ncell
        b_fwd = 8.0 * n_img + 4.0 * npx
        b_adj = 16.0 * n_adj + 4.0 * npx + 16.0 * ncell_img
        t_adj = max(t_fa - t_f, 1e-9)
        out["image"] = {"rays_per_s_fwd": 65536 / t_f, "rays_per_s_fwd_adj": 65536 / t_fa,
                        "ms_fwd": t_f * 1e3, "ms_fwd_adj": t_fa * 1e3,
                        "config": "256x256 px, 1.3 mm, incl 60, 2 arcsec field at 140 pc",
                        "ray_tile": list(tile),
                        "crossings_per_ray": n_img / npx,
                        "roofline_fwd": {"kernel": "k_image (+ k_cam_prep, k_stars in the "
                                         "timing)", "bound": "hbm",
                                         "achieved": b_fwd / t_f / 1e9, "peak": HBM_PEAK_GBS,
                                         "unit": "GB/s", "frac": b_fwd / t_f / 1e9 / HBM_PEAK_GBS,
                                         "b_alg": b_fwd,
                                         "note": "per-cell data (8 B x cells) stay in L2; "
                                                 "the kernel is bound by VALU issue per "
                                                 "crossing, not by HBM"},
                        "roofline_adj": {"kernel": "k_image_adj (+ k_cam_prep, k_cam_post)",
                                         "bound": "hbm", "achieved": b_adj / t_adj / 1e9,
                                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                         "frac": b_adj / t_adj / 1e9 / HBM_PEAK_GBS,
                                         "b_alg": b_adj}}
        # with the scattering source, as RADMC-3D images when the dust scatters:
        # J from the monochromatic Monte Carlo, then image + adjoint
        nps = int(a.nphot_scat)
        if nps > 0:
            dm.scattering_source(1300.0, nps, T, seed=1)      # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                J, sst = dm.scattering_source(1300.0, nps, T, seed=0)
            torch.cuda.synchronize()
            t_s = (time.perf_counter() - t0) / 3
            cam_s = dm.camera(256, pix_cm, 1300.0, incl=60.0, tile=tile, jscat=J)
            t0 = time.perf_counter()
            for _ in range(n_rep):
                dm.image(cam_s, T)
                dm.image_adjoint(cam_s, T, None, dimg)
            torch.cuda.synchronize()
            t_sa = (time.perf_counter() - t0) / n_rep
            out["image"]["scattering"] = {
                "nphot_scat": nps, "source_ms": t_s * 1e3, "source_packets_per_s": nps / t_s,
                "source_kernel_ms": sst["kernel_ms"], "ms_fwd_adj": t_sa * 1e3,
                "rays_per_s_fwd_adj_incl_source": 65536 / (t_s + t_sa),
                "n_scat": sst["n_scat"], "n_abs": sst["n_abs"], "n_escape": sst["n_escape"]}
        # visibilities of that image at the workload's 5e4 (u,v) points and at
        # 1e6 (a large interferometric data set): exact DFT and galario's method
        img = dm.image(cam, T)
        dxy = 2.0 / 256 * inputs.arcsec
        vis_out = {}
        for K in (50000, 1000000):
            u, v = configs.uv_points(K=K, lam_m=1.3e-3, seed=0)
            u = torch.tensor(u, device="cuda")
            v = torch.tensor(v, device="cuda")
            row = {}
            for name, fn in (("dft", rt.vis_dft), ("grid", rt.vis_grid)):
                fn(img, dxy, u, v)
                torch.cuda.synchronize()
                reps = 5
                t0 = time.perf_counter()
                for _ in range(reps):
                    fn(img, dxy, u, v)
                torch.cuda.synchronize()
                row[name + "_ms"] = (time.perf_counter() - t0) / reps * 1e3
            # its adjoint (the image gradient of chi^2 from the visibilities')
            gre = torch.ones(K, dtype=torch.float32, device="cuda")
            gim = torch.full((K,), 0.5, dtype=torch.float32, device="cuda")
            rt.vis_dft_adjoint(gre, gim, 256, 256, dxy, u, v)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                rt.vis_dft_adjoint(gre, gim, 256, 256, dxy, u, v)
            torch.cuda.synchronize()
            row["dft_adj_ms"] = (time.perf_counter() - t0) / reps * 1e3
            # exact DFT of a real image: a real pixel times a complex phasor
            # per pixel and point, 2 FMAs = 4 FP32 flops; the adjoint likewise
            # (the real part of a complex product per pixel and point)
            fl = 4.0 * 65536 * K
            row["dft_tflops"] = fl / (row["dft_ms"] * 1e-3) / 1e12
            row["dft_frac_fp32_vector"] = row["dft_tflops"] / FP32_VECTOR_TFLOPS
            row["dft_adj_tflops"] = fl / (row["dft_adj_ms"] * 1e-3) / 1e12
            row["dft_adj_frac_fp32_vector"] = row["dft_adj_tflops"] / FP32_VECTOR_TFLOPS
            vis_out["K=%d" % K] = row
        out["visibilities"] = dict(vis_out, config="256x256 image above, 1.3 mm, (u,v) of "
                                   "15-2000 m baselines",
                                   peak_note="FP32 vector peak without packed FP32 "
                                             "(256 CU x 4 SIMD x 16 lanes x 2 flop x 2.4 GHz "
                                             "= 78.6 TFLOP/s; the packed / MFMA figure is "
                                             "157.3); packed FP32 is off in this build")
        log("image legs done")
    if not a.no_cpu:
        out["accuracy"] = accuracy_vs_oracle(dm, arr, int(a.acc_nphot), 0, a.max_events, a.mrw,
                                             cpu["threads"])
        log("accuracy vs oracle at %g packets: median %.2e p99 %.2e"
            % (a.acc_nphot, out["accuracy"]["median_dT_over_T"],
               out["accuracy"]["p99_dT_over_T"]))
        cb = cpu_baseline(arr, a.max_events, a.cpu_seconds, a.mrw, cpu)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
        out["gpu_over_cpu_steps"] = roof["steps_per_s"] / cb["steps_per_s"]
        log("cpu baseline %.3e packets/s on %d threads" % (cb["value"], cb["cores"]))

    # the workload's four seeds in one batched launch, timed once beside the
    # steps (a step budget permitting): each run bit-identical to its plain run
    if a.seeds > 1 and (a.side_budget_s - elapsed()) > 4.5 * max(last or 0.0, 1.0):
        N = int(a.seeds)
        models = [dm] + [rt.DeviceModel(arr) for _ in range(N - 1)]
        batch = rt.Batch(models)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bst = batch.thermal([nphot] * N, list(range(N)), **opts)
        bT = [m.temperature() for m in models]
        t_b = time.perf_counter() - t0
        out["batched_seeds"] = {
            "runs": N, "seeds": list(range(N)), "s": t_b, "packets_per_s": N * nphot / t_b,
            "kernel_ms": bst[0]["kernel_ms"],
            "seed0_bitwise_equal_to_plain_run": bool(np.array_equal(bT[0], temps[0]))}
        log("batched launch of %d seeds: %.3e packets/s" % (N, N * nphot / t_b))
    out["wall_s_in_process"] = elapsed()
    print(json.dumps(out), flush=True)


def _with_T(model, T, dm):
    nr, nt, np_ = dm.nr, dm.nt, dm.np
    model.grid.temperature = [np.transpose(T[s].reshape((np_, nt, nr)), axes=(2, 1, 0))
                              for s in range(dm.nspec)]
    return model.grid


if __name__ == "__main__":
    main()
