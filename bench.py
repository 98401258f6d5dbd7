"""Benchmark of the MI355X DEM step path.

One step = one DEM time step (bin, detect, contact forces, integration,
geometry update) over the whole model, with the model resident in HBM.
Default workload: 2,000,000 polydisperse spheres (lognormal radius, median
2 mm, sigma 0.25, clipped 1-4 mm) poured into a conical hopper of ~5,000
static triangles, granular law k_n = 1e6 N/m, damping ratio 0.1, mu = 0.3,
dt = 2e-6 s, FP64.

Prints one JSON line (see the bench contract in the repository docs).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s plain FP64 vs batched mixed; p50/p99 sweep-request latency"
HBM_PEAK_GBS = 8000.0


def host_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def exact_duration(k, dt):
    """Duration for which dem() takes exactly k steps of size dt."""
    return (k - 0.5) * dt


def cpu_baseline(model, dt, steps, threads):
    """Time the CPU oracle (C++/OpenMP restatement of the reference step) on a
    bounded sample of the same workload: one untimed step, then `steps` timed."""
    from oracle import orc
    from parmec_amd import pmx
    o = orc.Oracle(threads)
    m = model.copy()
    o.upload(m)
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    o.dem(exact_duration(1, dt), dt, f)
    t0 = time.perf_counter()
    o.dem(exact_duration(steps, dt), dt, f)
    sec = time.perf_counter() - t0
    st = o.stats()
    o.close()
    return model.parnum * st["steps"] / sec, st["steps"], sec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=2000000)
    ap.add_argument("--triangles", type=int, default=5000)
    ap.add_argument("--dt", type=float, default=2e-6)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true", help="skip the CPU leg and JSON extras")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from parmec_amd import pmx, workloads

    t_setup = time.perf_counter()
    model = workloads.hopper(n=args.particles, ntri=args.triangles, seed=1)
    print(f"[bench] workload built: {model.parnum} spheres, {model.trinum} triangles "
          f"({time.perf_counter() - t_setup:.1f} s)", file=sys.stderr, flush=True)

    eng = pmx.Engine()
    m = model.copy()
    eng.upload(m)
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    eng.dem(exact_duration(args.warmup, args.dt), args.dt, f)
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    lib = eng.lib
    lib.pmx_set_profiling(eng.h, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.dem(exact_duration(args.steps, args.dt), args.dt, f)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    st = eng.stats()
    assert st["steps"] == args.steps, (st, args.steps)
    print(f"[bench] timed {args.steps} steps in {sec:.3f} s", file=sys.stderr, flush=True)

    import ctypes as C
    cnt = C.c_int(0)
    arr = (pmx.pmx_kernel_time * 16)() if hasattr(pmx, "pmx_kernel_time") else None
    lib.pmx_get_profile.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    lib.pmx_get_profile(eng.h, C.cast(arr, C.c_void_p), 16, C.byref(cnt))
    kernels = {}
    for k in range(cnt.value):
        e = arr[k]
        if e.launches:
            kernels[e.name.decode()] = {"launches": e.launches, "ms": e.ms, "bytes": e.bytes}
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    d = kernels[dom]
    avg_s = d["ms"] / d["launches"] * 1e-3
    bytes_per_launch = d["bytes"] / d["launches"]
    achieved = bytes_per_launch / avg_s / 1e9
    contacts = eng.contacts()
    pairs_per_particle = 2.0 * len(contacts) / model.parnum

    value = model.parnum * args.steps / sec
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"hopper: {model.parnum} polydisperse spheres (lognormal r, median 2 mm, sigma 0.25, "
                        f"clipped 1-4 mm) poured into a 60-degree conical hopper of {model.trinum} static "
                        f"triangles, granular kn=1e6 zeta=0.1 mu=0.3, dt={args.dt}, plain FP64",
            "particles": model.parnum,
            "triangles": model.trinum,
            "dt": args.dt,
            "precision": "fp64",
            "contact_pairs_per_particle": pairs_per_particle,
            "tree_rebuilds": st["rebuilds"],
            "leaves": st["leaves"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_us": avg_s * 1e6,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "kernels_ms_per_step": {k: v["ms"] / args.steps for k, v in kernels.items()},
    }
    eng.close()

    if not args.no_cpu and not args.profile_only:
        cpu_model, ncpu = host_info()
        threads = args.cpu_threads or min(ncpu or 1, 16)
        print(f"[bench] CPU oracle sample: {args.cpu_steps} steps on {threads} threads", file=sys.stderr, flush=True)
        v, steps, csec = cpu_baseline(model, args.dt, args.cpu_steps, threads)
        out["cpu_baseline"] = {
            "value": v,
            "unit": "particle-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{steps} steps of the same {model.parnum}-sphere hopper after one untimed step "
                      f"({csec:.1f} s), C++/OpenMP restatement of the reference step on {cpu_model}",
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
