"""Benchmark of the MI355X DEM step path.

Headline (value): the batched acceleration sweep in mixed precision.  512
independent instances of the 10,000-sphere box set-up (sphere count uniform in
[1k, 64k], base excitation g(t) = (A sin 2 pi f t, 0, -9.81) with A in
[0.1, 2] g and f in [1, 20] Hz, seed = instance number), dt = 1e-5 s, stepped
together in one batch (one fused launch per stage, the time step replayed as a
HIP graph), FP32 contact law / FP64 integration.  The synthetic input is the
sweep's state after --settle steps (default 3000, run on the GPU before
timing: the spheres have fallen into contact), then --warmup steps, then
--steps timed steps, all in one dem call as a long run takes them (the
call's own set-up, the partitioning tree of every instance and the step
graph, falls at its start).  The timed steps are measured on the device
timeline, from an event recorded on the library's stream before the first
of them to one after the last (pmx_set_window; a step repeated after a leaf
overflow re-records its event, so a rebuild inside the window is counted),
with the state and all inputs resident in HBM; the host clock of the whole
call is reported beside it.

Beside it, in the same JSON line:
  plain_fp64  -- the drop-in FP64 path stepping sweep instances one at a time
                 (a sample of instances over the same window), the
                 comparison the batched mixed form is measured against;
  dense       -- the same 512 instance sizes as confined dense packs (every
                 sphere pressed against its six neighbours: ~3 contact pairs
                 per sphere, the regime of a settled pack), so that the
                 contact stage is timed where it matters;
  hopper      -- 2,000,000 polydisperse spheres in a 5,000-triangle hopper,
                 FP64 and mixed, dt = 2e-6 s;
  serving     -- a Poisson stream (20 requests/s) of independent simulations
                 of uneven size (70 % sphere boxes, 30 % spring lattices,
                 lognormal size, median 20k, 500-20,000 steps) submitted in
                 real time to the resident request loop: p50/p99 latency and
                 particle-steps/s, batched mixed and batched FP64 beside the
                 plain FP64 path serving one request at a time;
  roofline    -- the dominant kernel of the timed window: algorithmic bytes
                 per launch (the bytes its branches read and write) over its
                 mean duration from HIP event record nodes around every 8th
                 step's kernels on the library's stream;
  step_roofline -- the whole step: particle-steps/s x the compulsory bytes of
                 a fused step per particle (384 + 40 P mixed, 384 + 64 P FP64,
                 P = contact pairs per particle) over 8 TB/s;
  cpu_baseline -- the C++/OpenMP restatement of the reference step (the
                 oracle) on a bounded sample of the same sweep state, on all
                 of the job's host cores and on one core.

One step = one DEM time step (bin, detect, contact forces, integration,
geometry update) over every instance.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s plain FP64 vs batched mixed; p50/p99 sweep-request latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def kernel_table(pmx, eng):
    arr = (pmx.pmx_kernel_time * 32)()
    cnt = C.c_int(0)
    eng.lib.pmx_get_profile.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    eng.lib.pmx_get_profile(eng.h, C.cast(arr, C.c_void_p), 32, C.byref(cnt))
    out = {}
    for k in range(cnt.value):
        e = arr[k]
        if e.launches:
            out[e.name.decode()] = {"launches": int(e.launches), "ms": float(e.ms), "bytes": float(e.bytes)}
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_sweep512_mixed.json")


def measured_traffic(kernel, instances, spheres, precision, window):
    """HBM bytes per launch of a step kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by the copy-probe factors), when
    they were taken on this workload and this timed window; None otherwise."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    if (t.get("instances"), t.get("spheres"), t.get("precision")) != (instances, spheres, precision):
        return None, None
    if list(t.get("window", [])) != list(window):
        return None, (f"{os.path.relpath(TRAFFIC_FILE, ROOT)} holds steps {t.get('window')}, not the timed "
                      f"window {list(window)}")
    k = t["kernels"].get(kernel)
    if not k:
        return None, None
    b = 1024.0 * (t["fetch_factor"] * k["fetch_kib"] + t["write_factor"] * k["write_kib"])
    src = (f"{os.path.relpath(TRAFFIC_FILE, ROOT)}: rocprofv3 FETCH_SIZE x {t['fetch_factor']} + WRITE_SIZE x "
           f"{t['write_factor']} per dispatch, steps {t['window'][0]}-{t['window'][1]}")
    return b, src


def roofline(kernels, extra_bytes, instances, spheres, precision, window, graph_timed):
    """The step kernel with the largest total time (kernels with algorithmic
    bytes; the tree rebuild is reported in kernels_ms_per_launch).  Launch
    durations come from HIP event pairs on the context's stream around every
    launch of the profiled steps: event record nodes of the replayed step
    graph (graph_timed), else eager steps.  extra_bytes: bytes per launch the
    library's count leaves out (the force and torque reads of particles
    whose row is not empty, by kernel)."""
    dom = max((k for k in kernels if kernels[k]["bytes"] > 0), key=lambda k: kernels[k]["ms"])
    d = kernels[dom]
    avg_s = d["ms"] / d["launches"] * 1e-3
    per_launch = d["bytes"] / d["launches"] + extra_bytes.get(dom, 0.0)
    achieved = per_launch / avg_s / 1e9
    traffic, src = measured_traffic(dom, instances, spheres, precision, window)
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": src,
            "avg_launch_us": avg_s * 1e6, "algorithmic_bytes_per_launch": per_launch,
            "sampled_launches": d["launches"],
            "timing": "event record nodes in the replayed step graph" if graph_timed else "eager steps with event pairs"}


def step_roofline(value, pairs_per_particle, mixed):
    """The whole step against the HBM roof: the compulsory bytes of a fused
    step per particle, 384 + 40 P (mixed: FP32 pair state) or 384 + 64 P."""
    per = 384.0 + (40.0 if mixed else 64.0) * pairs_per_particle
    gbs = value * per / 1e9
    return {"bytes_per_particle_step": per, "pairs_per_particle": pairs_per_particle, "achieved": gbs,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}


def per_step(kernels, nsteps):
    """Kernel time per step: every per-step kernel has one profiled scope per
    profiled step (the library times every 8th step); rebuilds are timed
    whenever they happen, so their total is spread over all nsteps steps."""
    return {k: (v["ms"] / v["launches"] if k != "tree_build" else v["ms"] / nsteps) for k, v in kernels.items()}


def sweep_forcings(pmx, workloads, insts, nsteps, dt, t0):
    return [pmx.forcing(nsteps, gtab=workloads.sweep_forcing(a, f, nsteps, dt, t0)) for (_, a, f) in insts]


def plain_fp64_sample(pmx, workloads, insts, sample, settle, warm, steps, dt):
    """The drop-in FP64 path, one instance at a time, over the same window,
    timed the same way (each sampled instance stepped alone from rest in one
    call, the window's steps timed on the device)."""
    import torch
    tot_ps, tot_s, ids = 0, 0.0, []
    for i in sample:
        m, a, f = insts[i]
        m = m.copy()
        e = pmx.Engine()
        e.upload(m)
        n = settle + warm + steps
        e.lib.pmx_set_window.argtypes = [C.c_void_p, C.c_int, C.c_int]
        e.lib.pmx_set_window(e.h, settle + warm, n - 1)
        fs = pmx.forcing(n, gtab=workloads.sweep_forcing(a, f, n, dt, 0.0))
        e.dem(exact_duration(n, dt), dt, fs)
        torch.cuda.synchronize()
        st = e.stats()
        e.close()
        assert st["steps"] == n and st["window_ms"] > 0
        sec = st["window_ms"] / 1e3
        tot_ps += m.parnum * steps
        tot_s += sec
        ids.append(i)
    return tot_ps / tot_s, ids


def dense_leg(pmx, workloads, args, dt):
    """The sweep's 512 instance sizes as confined dense packs (about 3
    contact pairs per sphere from the first step), batched, mixed precision:
    the contact stage timed in the regime of a settled pack."""
    import torch
    insts = [workloads.dense_instance(i, args.nmin, args.nmax) for i in range(args.instances)]
    models = [x[0] for x in insts]
    ntot = sum(m.parnum for m in models)
    eng = pmx.Engine(precision=pmx.PMX_MIXED)
    eng.upload_batch(models)
    W, K = args.dense_warmup, args.dense_steps
    # warm-up and timed steps in one call, the timed steps measured on the
    # device and profiled (as the sweep's window)
    eng.lib.pmx_set_profiling(eng.h, 1)
    eng.lib.pmx_set_window.argtypes = [C.c_void_p, C.c_int, C.c_int]
    eng.lib.pmx_set_window(eng.h, W, W + K - 1)
    eng.dem_batch(exact_duration(W + K, dt), dt, sweep_forcings(pmx, workloads, insts, W + K, dt, 0.0))
    torch.cuda.synchronize()
    st = eng.stats()
    sec = st["window_ms"] / 1e3
    kernels = kernel_table(pmx, eng)
    eng.close()
    value = ntot * K / sec
    busy = st["busy"]
    pairs = pair_count(st, ntot)
    log(f"dense packs: {value:.3e} particle-steps/s, {pairs:.2f} contact pairs per particle, rebuilds {st['rebuilds']}")
    return {"value": value, "unit": "particle-steps/s", "instances": len(models), "spheres": ntot,
            "window": [W, W + K], "ms_per_step": 1e3 * sec / K, "contact_pairs_per_particle": pairs,
            "precision": "mixed",
            "workload": "the sweep's instance sizes, amplitudes and frequencies as confined dense packs "
                        "(lattice spacing 1.98 r, jitter 0.004 r, walls overlapping the outer spheres by 0.02 r)",
            "step_roofline": step_roofline(value, pairs, True),
            "roofline": roofline(kernels, {"dynamics": 40.0 * busy}, len(models), ntot, "dense", [W, W + K],
                                 st["graph_timed"]),
            "kernels_ms_per_step": per_step(kernels, K)}


def pair_count(st, ntot):
    """Contact pairs per particle from the row entries (sphere-sphere pairs
    are held in both spheres' rows; triangle contacts once).  The split is
    not known from the totals, so pairs are counted as half the entries: a
    lower bound of P (exact without triangle contacts)."""
    return 0.5 * st["pairs"] / ntot


def hopper_leg(pmx, workloads, n, steps, dt):
    import torch
    model = workloads.hopper(n=n, seed=1)
    out = {"spheres": model.parnum, "triangles": model.trinum, "dt": dt, "steps": steps}
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    for prec, name in ((pmx.PMX_FP64, "fp64"), (pmx.PMX_MIXED, "mixed")):
        e = pmx.Engine(precision=prec)
        m = model.copy()
        e.upload(m)
        e.dem(exact_duration(20, dt), dt, f)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.dem(exact_duration(steps, dt), dt, f)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        out[name] = model.parnum * steps / sec
        e.close()
        log(f"hopper {name}: {out[name]:.3e} particle-steps/s")
    return out


def serve_leg(pmx, workloads, count, rate, seed):
    """A Poisson stream of independent requests (boxes of spheres and spring
    lattices of uneven size and length) submitted in real time to the
    resident request loop: batched in mixed precision and in FP64, and the
    plain FP64 path serving them one at a time.  Latency = result on the host
    - submission."""
    import numpy as np
    stream = workloads.serve_stream(count, rate=rate, seed=seed)
    out = {"requests": count, "rate_per_s": rate, "seed": seed,
           "spheres": int(sum(r[1].ellnum for r in stream)), "bodies": int(sum(r[1].parnum for r in stream)),
           "particle_steps": float(sum(r[1].parnum * int(round(r[2] / r[3] + 0.5)) for r in stream))}
    # batched in both precisions (batching and precision separated), the
    # continuous form (requests join running batches of their kind), and the
    # plain path: one request at a time, FP64
    batched = dict(gather_s=0.25, max_active=4, max_batch=16, longest_first=True)
    forms = (("batched_mixed", pmx.PMX_MIXED, batched),
             ("batched_fp64", pmx.PMX_FP64, batched),
             ("continuous_mixed", pmx.PMX_MIXED, dict(gather_s=0.05, max_active=4, max_batch=16, repack_s=0.25,
                                                      longest_first=True)),
             ("plain_fp64", pmx.PMX_FP64, dict(max_batch=1, max_active=1)))
    for name, prec, kw in forms:
        srv = pmx.Server(precision=prec, **kw)
        ids = []
        t0 = time.perf_counter()
        for (ta, m, d, h, f) in stream:
            wait = ta - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            ids.append(srv.submit(m, d, h, f))
        infos = [srv.wait(i) for i in ids]
        st = srv.stats()
        srv.close()
        lat = np.array([i["done_s"] - i["submit_s"] for i in infos])
        ps = float(sum(i["particles"] * i["steps"] for i in infos))
        span = max(i["done_s"] for i in infos) - min(i["submit_s"] for i in infos)
        out[name] = {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                     "value": ps / span, "unit": "particle-steps/s", "span_s": span, "batches": st["batches"]}
        log(f"serve {name}: p50 {out[name]['p50_s']:.2f} s, p99 {out[name]['p99_s']:.2f} s, "
            f"{out[name]['value']:.3e} particle-steps/s over {span:.1f} s, {st['batches']} batches")
    for name in ("batched_mixed", "batched_fp64", "continuous_mixed"):
        out[name]["speedup_vs_plain_fp64"] = out[name]["value"] / out["plain_fp64"]["value"]
    out["policy"] = ("batched: one kind of request per batch (spheres, or bodies joined by springs), gathered over "
                     "0.25 s windows, at most 16 requests per batch and 4 batches at once, the queued requests with "
                     "the most steps first; "
                     "continuous: a running batch of a kind (at most 16 requests, 4 batches at once) pauses at "
                     "most every 0.25 s to take in queued requests of its kind, carrying its unfinished requests "
                     "over with their contact rows and tree snapshots")
    return out


def cpu_leg(pmx, models, forcings_fn, ids, steps, threads, dt):
    """Oracle on a bounded sample: the sample instances' current states,
    stepped on the host's cores."""
    from oracle import orc
    tot_ps, tot_s = 0, 0.0
    for i in ids:
        o = orc.Oracle(threads)
        m = models[i].copy()
        o.upload(m)
        t0 = time.perf_counter()
        o.dem(exact_duration(steps, dt), dt, forcings_fn(i, steps))
        tot_s += time.perf_counter() - t0
        tot_ps += m.parnum * o.stats()["steps"]
        o.close()
    return tot_ps / tot_s, tot_ps, tot_s


def host_cpu():
    """nproc, the CPUs this process may run on, model name and SMT state of
    the host."""
    model, ncpu = host_info()
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = ncpu
    smt = None
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = f.read().strip() == "1"
    except OSError:
        pass
    return {"model": model, "nproc": ncpu, "allowed": allowed, "smt_active": smt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed steps of the sweep window")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps after the settled state, in the timed "
                    "steps' dem call (the clock starts at a pause after them)")
    ap.add_argument("--settle", type=int, default=3000,
                    help="steps run before the warmup to form the input state (the spheres fall into contact)")
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--nmin", type=int, default=1000)
    ap.add_argument("--nmax", type=int, default=64000)
    ap.add_argument("--dt", type=float, default=1e-5)
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed")
    ap.add_argument("--plain-sample", type=int, default=8, help="instances timed on the plain FP64 path")
    ap.add_argument("--dense-steps", type=int, default=200, help="timed steps of the dense-pack leg (0: skip)")
    ap.add_argument("--dense-warmup", type=int, default=20)
    ap.add_argument("--hopper", type=int, default=2000000, help="hopper spheres (0: skip)")
    ap.add_argument("--hopper-steps", type=int, default=200)
    ap.add_argument("--serve", type=int, default=200, help="requests of the serving stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--cpu-instances", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--cpu-threads", type=int, default=0, help="all-core leg threads (0: the CPUs this job may use, "
                                                             "at most 16, the host share of one GPU)")
    ap.add_argument("--cpu-single-instances", type=int, default=4, help="instances of the one-core leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--full-run", action="store_true",
                    help="instead of a window: the whole 10,000-step sweep from rest, end to end")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run an elementwise kernel reading and writing MIB MiB 10 times "
                         "(calibrates the HBM byte counters) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding of the path is a design, not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe > 0:
        # a kernel, not a DMA copy (which counters do not see): bitwise_not
        # reads n bytes and writes n bytes per dispatch
        n = args.copy_probe << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.bitwise_not(a, out=b)
        e0.record()
        for _ in range(9):
            torch.bitwise_not(a, out=b)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 9
        # the practical streaming rate of one elementwise kernel (read n, write n)
        print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n, "dispatches": 10,
                          "ms_per_dispatch": ms, "gb_per_s": 2.0 * n / ms / 1e6}))
        return

    from parmec_amd import pmx, workloads

    dt = args.dt
    if args.full_run:
        S0, W, K = 0, 0, 10000
    else:
        S0, W, K = args.settle, args.warmup, args.steps
    start = S0 + W
    t_setup = time.perf_counter()
    insts = [workloads.sweep_instance(i, args.nmin, args.nmax) for i in range(args.instances)]
    models = [x[0] for x in insts]
    ntot = sum(m.parnum for m in models)
    log(f"sweep built: {len(models)} instances, {ntot} spheres ({time.perf_counter() - t_setup:.1f} s)")

    prec = pmx.PMX_MIXED if args.precision == "mixed" else pmx.PMX_FP64
    eng = pmx.Engine(precision=prec)
    eng.upload_batch(models)
    done, rebuilds, grown = 0, 0, 0
    call_s = None
    eng.lib.pmx_set_window.argtypes = [C.c_void_p, C.c_int, C.c_int]
    if not args.full_run:
        # settling, warm-up and timed steps in one dem call, as a long run
        # takes them (the device busy throughout, the call's own set-up at
        # its start); the timed steps are measured on the device timeline
        # (events before step `start` and after the last step) and only
        # their kernels are profiled
        eng.lib.pmx_set_profiling(eng.h, 1)
        n = start + K
        eng.lib.pmx_set_window(eng.h, start, start + K - 1)
        fcs = sweep_forcings(pmx, workloads, insts, n, dt, 0.0)
        log(f"one dem call of {n} steps, steps {start}-{start + K - 1} timed on the device")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.dem_batch(exact_duration(n, dt), dt, fcs)
        torch.cuda.synchronize()
        call_s = time.perf_counter() - t0
        st = eng.stats()
        assert st["steps"] == n and st["window_ms"] > 0, (st, n)
        rebuilds, grown, done = st["rebuilds"], st["grown"], K
    else:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    while done < K:
        # the full run in calls of 1000 steps (forcing tables of that length)
        n = min(1000, K - done)
        eng.dem_batch(exact_duration(n, dt), dt, sweep_forcings(pmx, workloads, insts, n, dt, (start + done) * dt))
        st = eng.stats()
        assert st["steps"] == n, (st, n)
        rebuilds += st["rebuilds"]
        grown += st["grown"]
        done += n
        log(f"full run {done}/{K} steps, {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0 if args.full_run else st["window_ms"] / 1e3
    kernels = kernel_table(pmx, eng) if not args.full_run else {}
    eng.lib.pmx_set_profiling(eng.h, 0)
    value = ntot * K / sec
    busy = st["busy"]
    pairs = pair_count(st, ntot)
    log(f"timed {K} steps in {sec:.3f} s: {value:.3e} particle-steps/s, pairs {st['pairs']}, rebuilds {rebuilds}")

    state = None
    if not args.no_cpu:
        state = [m.copy() for m in models]
        eng.download_batch(state)
    eng.close()

    stride = max(1, args.instances // max(1, args.plain_sample))
    sample = list(range(0, args.instances, stride))[: args.plain_sample]
    plain, plain_ids = None, []
    if sample and not args.full_run:
        plain, plain_ids = plain_fp64_sample(pmx, workloads, insts, sample, S0, W, min(K, 1000), dt)
        log(f"plain fp64 (one instance at a time): {plain:.3e} particle-steps/s")

    dense = dense_leg(pmx, workloads, args, dt) if args.dense_steps > 0 and not args.full_run else None
    hopper = hopper_leg(pmx, workloads, args.hopper, args.hopper_steps, 2e-6) if args.hopper > 0 else None
    serving = serve_leg(pmx, workloads, args.serve, args.serve_rate, 0) if args.serve > 0 else None

    cpu = None
    if not args.no_cpu:
        host = host_cpu()
        threads = args.cpu_threads or max(1, min(16, host["allowed"] or 1))
        ids = list(range(0, args.instances, max(1, args.instances // args.cpu_instances)))[: args.cpu_instances]
        t_at = (start + K) * dt

        def ffn(i, n):
            _, a, f = insts[i]
            return pmx.forcing(n, gtab=workloads.sweep_forcing(a, f, n, dt, t_at))
        rate, ps, s = cpu_leg(pmx, state, ffn, ids, args.cpu_steps, threads, dt)
        ids1 = ids[: args.cpu_single_instances]
        rate1, ps1, s1 = cpu_leg(pmx, state, ffn, ids1, args.cpu_steps, 1, dt)
        cpu = {"value": rate, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "single_core": rate1,
               "host": host,
               "build": "g++ -O3 -march=x86-64-v3 -ffp-contract=off, OpenMP (oracle/Makefile)",
               "sample": f"{len(ids)} sweep instances ({ps // args.cpu_steps} spheres) at their state after step "
                         f"{start + K}, {args.cpu_steps} steps each on {threads} threads ({s:.1f} s), and "
                         f"{len(ids1)} of them on one thread ({s1:.1f} s); the C++/OpenMP restatement of the "
                         f"reference step (the oracle), not the ISPC build"}
        log(f"cpu baseline: {rate:.3e} particle-steps/s on {threads} threads, {rate1:.3e} on one")

    window = [start, start + K]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * sec / K,
        "timing": ({"how": "device events around the timed steps inside one dem call (pmx_set_window)",
                    "call_seconds": call_s, "call_steps": start + K,
                    "call_includes": "settling, warm-up and timed steps, the call's tree builds and graph "
                                     "capture, its host step loop"}
                   if call_s is not None else {"how": "host clock around the calls of the full run"}),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 contact law, f64 state" if prec == pmx.PMX_MIXED else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"sweep: {len(models)} instances of the box set-up (r = 5 mm spheres on a jittered "
                        f"lattice in a 12-triangle box, kn = 1e6, zeta = 0.1, mu = 0.3), {args.nmin}-{args.nmax} "
                        f"spheres each (seeded), base excitation 0.1-2 g at 1-20 Hz, dt = {dt}, batched, "
                        f"{args.precision}; " + ("the whole 10,000-step run from rest" if args.full_run else
                                                 f"input = the state after {S0} steps, timed steps "
                                                 f"{start}-{start + K} of the 10,000-step run (one dem call of "
                                                 f"steps 0-{start + K}; the timed steps on the device timeline)"),
            "instances": len(models),
            "spheres": ntot,
            "dt": dt,
            "precision": args.precision,
            "window": window,
            "contact_pairs_per_particle": pairs,
            "spheres_in_contact": st["busy"],
            "tree_rebuilds": rebuilds,
            "capacity_regrowths": grown,
            "leaves": st["leaves"],
            "block_shapes": {"detect_waves_per_block": st["det_waves"], "detect_autotuned": True,
                             "rows_threads_per_block": st["rows_block"]},
        },
        "plain_fp64": {"value": plain, "unit": "particle-steps/s", "instances_sampled": plain_ids,
                       "how": "drop-in FP64 path, each sampled instance alone from rest, timed over the same window "
                              "the same way (clock started after the warm-up steps of the timed call)"},
        "speedup_vs_plain_fp64": value / plain if plain else None,
        "step_roofline": step_roofline(value, pairs, prec == pmx.PMX_MIXED),
        "dense": dense,
        "hopper": hopper,
        "serving": serving,
        "roofline": roofline(kernels, {"dynamics": 40.0 * busy}, len(models), ntot, args.precision, window,
                             st["graph_timed"]) if kernels else None,
        "kernels_ms_per_launch": {k: v["ms"] / v["launches"] for k, v in kernels.items()},
        "kernels_ms_per_step": per_step(kernels, K) if kernels else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
