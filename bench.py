"""Throughput of the sLLG ensemble path on one MI355X.

Default workload (the headline line): the 65,536-particle AC-field ensemble —
65,536 independent particles, radii lognormal (median 10 nm, sigma 0.1),
K = 23 kJ/m^3, easy axes uniform on the sphere, m0 along the easy axis,
alpha = 0.1, T = 300 K, sinusoidal z field 20 mT at 300 kHz, Heun, FP32,
dt = 1e-14 s, run length 5 field cycles (1.67e9 steps) with 2,000 samples per
cycle reduced on the device.  Inputs are synthetic and resident in HBM
before timing starts.

Other workloads (--workload):
  clusters  4,096 interacting clusters, N uniform in [2, 64], 10% volume
            fraction, centres >= 2.2 radii apart, radius 10 nm, zero field,
            Heun FP32, dt = 1e-14 s, 2e5 steps.
  implicit  the 65,536-particle AC-field ensemble with the implicit midpoint
            scheme at dt = 5e-13 s (FP64, per-particle Newton); the run ends
            after the timed blocks so the Newton counts can be reported.
  replicas  one 512-particle cluster x 4,096 thermal replicas, dipolar field
            as an MFMA product (BF16 three-term split).
  adjoint   16,384 AC-field particles: the recording forward pass (forward
            kernel, and on a second stream the adjoint maps of the Heun step
            between sample steps) and the reverse scan; gradient of a
            weighted sum of <M_z> samples with respect to (K, alpha, H,
            median r).
  implicit_clusters  the 4,096 clusters of 'clusters' with the implicit
            midpoint scheme (FP64, dt = 5e-13 s): per-particle 3 x 3 Newton
            solves with the cluster's dipolar field at every iterate and
            cluster-wide norms (the reference's default Model.simulate path).
  equilibrium  1,000 independent 12 nm particles along +z in zero field, Heun
            FP64, 1e5 steps of 1e-14 s (the plain CPU-runnable ensemble; its
            CPU baseline times whole 1e5-step runs of the oracle).
  dom       the discrete orientation model of the 65,536 AC-field particles
            (two-state master equation, adaptive Cash-Karp RK45 in FP64,
            tolerance 1e-12, 5 field cycles, 11 samples); one bench step is
            one whole run, timed by HIP events around its kernel; value in
            accepted RK45 particle-steps/s.

--grad-check runs the full differentiation workload (2 cycles at 300 kHz,
6.67e8 steps) and a finite-difference check of its gradients on a 30 MHz
field; --sar-run runs the whole AC-field workload for its SAR.

One bench "step" is one step block: BLOCK time steps of every particle, one
kernel launch.  value = particle-steps / wall time of the timed blocks.  The
roofline object divides the algorithmic work of one launch (FLOP counted
from the reference's formulas) by the average launch duration measured with
HIP events on the launch stream.  The CPU baseline times the FP64 oracle (a
reference-structured restatement of simulation::full_dynamics with the
reference's mt19937_64 Wiener source) on the host cores on a bounded sample
of the same workload.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--block B] [--workload W]
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MU0 = 1.25663706e-6
METRIC = "particle-steps/s (sLLG Heun, 1 GPU) at SAR err ≤0.5% vs FP64; plain vs batched"
FLOP_PER_PARTICLE_STEP = 278          # Heun, N = 1, no renorm
# implicit midpoint, N = 1 (lib/integrators.cpp:576-651 with Newton,
# lib/optimisation.cpp:81-149): per Newton iteration, and per step besides
# the iterations (Euler guess with its sde and Jacobians, tolerance, final
# update); counted on the oracle's arithmetic by oracle/flop_count.py under
# the convention that gives the Heun step its 278 (tests/test_flop_count.py)
FLOP_IMPLICIT_PER_NEWTON = 503
FLOP_IMPLICIT_PER_STEP = 365
FLOP_PER_PAIR_STEP = 48               # 24 FLOP per ordered pair, two field evaluations
# one accepted Cash-Karp step of the master equation, counted on the oracle's
# arithmetic (oracle/flop_count.py dom_flops())
FLOP_DOM_PER_STEP = 401
PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}   # MI355X vector peaks (published)
# HBM bytes of one default launch (65,536 particles x 131,072 steps), from
# rocprofv3 FETCH_SIZE (x2.000) + WRITE_SIZE (x1.000) scaled against a 1 GiB
# device copy; see profiles/heun_indep_pc_fp32_65536_ac_field.md.
TRAFFIC_BYTES_DEFAULT_LAUNCH = 9.26e6
# The metric's accuracy condition, recorded by full runs of the default
# workload (--sar-run fp32|fp64, 1.67e9 steps each; FP64 in three resumed
# sittings); see profiles/sar_fp32_vs_fp64_ac_field_65536.md.  Reported
# beside the default line, not measured by it.
SAR_ACCURACY_RECORDED = {
    "recorded_in": "profiles/sar_fp32_vs_fp64_ac_field_65536.md",
    "run": "65,536 particles, 5 cycles of 20 mT / 300 kHz, 1.67e9 steps, final-cycle energy",
    "sar_fp64_W_per_m3": [-1.45442e8, -1.48240e8], "sar_fp32_W_per_m3": [-1.45097e8, -1.48327e8],
    "noise_tags": [0, 1], "fp32_vs_fp64_rel": [2.38e-3, 5.86e-4], "within_0.5pct": True,
    "between_thermal_histories_rel_fp64": 1.89e-2}
DEFAULTS = {"ac_field": (65536, 131072), "clusters": (4096, 8192), "implicit": (65536, 256),
            "replicas": (4096, 64), "adjoint": (16384, 1048000), "dom": (65536, 1),
            "equilibrium": (1000, 4096), "implicit_clusters": (4096, 32)}


def ac_field_inputs(n, seed=2024):
    rng = np.random.default_rng(seed)
    r = 10e-9 * np.exp(0.1 * rng.normal(size=n))
    th = 2 * np.pi * rng.uniform(size=n)
    ph = np.arccos(1 - 2 * rng.uniform(size=n))
    axes = np.stack([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)], 1)
    return dict(member_offsets=np.arange(n + 1, dtype=np.int64), radius=r,
                anisotropy=np.full(n, 23e3), anisotropy_axis=axes,
                magnetisation_direction=axes.copy(), location=np.zeros((n, 3)),
                magnetisation=400e3, damping=0.1, temperature=300.0, field_shape='sine',
                field_amplitude=0.02 / MU0, field_frequency=3e5,
                seeds=rng.integers(0, 2 ** 31 - 1, size=n))


def cluster_inputs(n_clusters, seed=2025):
    from magpy_amd import geometry as G
    rng = np.random.default_rng(seed)
    sizes = rng.integers(2, 65, size=n_clusters)
    e = G.cluster_ensemble(sizes, 10e-9, 0.10, 2.2, rng)
    e.update(magnetisation=400e3, damping=0.1, temperature=300.0, field_shape='constant',
             field_amplitude=0.0, field_frequency=0.0,
             seeds=rng.integers(0, 2 ** 31 - 1, size=n_clusters))
    return e


def replica_inputs(n_replicas, n=512, seed=2026):
    from magpy_amd import geometry as G
    rng = np.random.default_rng(seed)
    loc = G.random_cluster(n, 10e-9, 0.15, 2.2, rng)
    axes = G.uniform_random_axes(n, rng)
    P = n * n_replicas
    return dict(member_offsets=np.arange(n_replicas + 1, dtype=np.int64) * n,
                radius=np.full(P, 10e-9), anisotropy=np.full(P, 23e3),
                anisotropy_axis=np.tile(axes, (n_replicas, 1)),
                magnetisation_direction=np.tile(axes, (n_replicas, 1)),
                location=np.tile(loc, (n_replicas, 1)), magnetisation=400e3, damping=0.1,
                temperature=300.0, field_shape='sine', field_amplitude=0.02 / MU0,
                field_frequency=3e5, seeds=rng.integers(0, 2 ** 31 - 1, size=n_replicas))


def equilibrium_inputs(n, seed=2023):
    """Independent 12 nm particles, easy axis and initial state along +z, no
    applied field: the plain FP64 thermal-equilibrium ensemble."""
    rng = np.random.default_rng(seed)
    z = np.tile([0.0, 0.0, 1.0], (n, 1))
    return dict(member_offsets=np.arange(n + 1, dtype=np.int64), radius=np.full(n, 12e-9),
                anisotropy=np.full(n, 23e3), anisotropy_axis=z, magnetisation_direction=z.copy(),
                location=np.zeros((n, 3)), magnetisation=400e3, damping=0.1, temperature=300.0,
                field_shape='constant', field_amplitude=0.0, field_frequency=0.0,
                seeds=rng.integers(0, 2 ** 31 - 1, size=n))


def workload(name, n, timed_steps):
    """(inputs, session kwargs, algorithmic FLOP per time step, precision, description)."""
    if name == "ac_field":
        inp = ac_field_inputs(n)
        kw = dict(time_step=1e-14, end_time=5 / 3e5, max_samples=2000 * 5 + 1,
                  interactions=False, use_implicit=False)
        return (inp, kw, float(FLOP_PER_PARTICLE_STEP * n), "fp32",
                "65,536-particle AC-field Heun ensemble (20 mT, 300 kHz, dt=1e-14 s)")
    if name == "equilibrium":
        inp = equilibrium_inputs(n)
        kw = dict(time_step=1e-14, end_time=1e5 * 1e-14, max_samples=101, interactions=False,
                  use_implicit=False)
        return (inp, kw, float(FLOP_PER_PARTICLE_STEP * n), "fp64",
                "1,000 independent 12 nm particles, zero field, Heun FP64 (dt=1e-14 s, "
                "1e5 steps: the plain thermal-equilibrium ensemble)")
    if name == "implicit_clusters":
        inp = cluster_inputs(n)
        kw = dict(time_step=5e-13, end_time=timed_steps * 5e-13, max_samples=2,
                  interactions=True, use_implicit=True)
        return (inp, kw, None, "fp64",
                "4,096 interacting clusters N~U{2..64}, 10% volume fraction, implicit midpoint "
                "FP64 with the dipolar field at every Newton iterate (dt=5e-13 s)")
    if name == "implicit":
        inp = ac_field_inputs(n)
        kw = dict(time_step=5e-13, end_time=timed_steps * 5e-13, max_samples=2,
                  interactions=False, use_implicit=True)
        return (inp, kw, None, "fp64",
                "65,536-particle AC-field ensemble, implicit midpoint FP64 (dt=5e-13 s)")
    if name == "clusters":
        inp = cluster_inputs(n)
        sizes = np.diff(inp["member_offsets"])
        flop = float(FLOP_PER_PARTICLE_STEP * sizes.sum()
                     + FLOP_PER_PAIR_STEP * np.sum(sizes * (sizes - 1)))
        kw = dict(time_step=1e-14, end_time=2e5 * 1e-14, max_samples=201, interactions=True,
                  use_implicit=False)
        return (inp, kw, flop, "fp32",
                "4,096 interacting clusters N~U{2..64}, 10% volume fraction, Heun FP32 "
                "(dt=1e-14 s, 2e5 steps)")
    if name == "replicas":
        inp = replica_inputs(n)
        n3 = 3 * 512
        flop = float(2 * 2 * n3 * n3 * n + FLOP_PER_PARTICLE_STEP * 512 * n)
        kw = dict(time_step=1e-14, end_time=1e5 * 1e-14, max_samples=101, interactions=True,
                  use_implicit=False)
        return (inp, kw, flop, "fp32",
                "one 512-particle cluster (15% volume fraction) x 4,096 thermal replicas, "
                "dipolar field as a BF16x3 MFMA product, Heun FP32 (20 mT, 300 kHz, "
                "dt=1e-14 s, 1e5 steps)")
    raise SystemExit("unknown workload " + name)


def host_cpu():
    """The host the CPU baseline runs on: its CPU model, logical CPUs, the
    CPUs this process may run on (affinity), a cgroup CPU quota if one is
    set, and the CPU share the job is given (OMP_NUM_THREADS, set by the
    job's environment).  The baseline uses every CPU of the affinity set
    within the quota and the share."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    share = os.environ.get("OMP_NUM_THREADS")
    share = int(share) if share and share.isdigit() and int(share) > 0 else None
    threads = affinity
    if quota is not None:
        threads = min(threads, max(1, int(quota)))
    if share is not None:
        threads = min(threads, share)
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": affinity,
            "cgroup_cpu_quota": quota, "job_cpu_share": share, "threads": threads}


def cpu_baseline(name, inp, seconds=12.0):
    """Time the oracle on the host cores (threads release the GIL in ctypes)."""
    from oracle import oracle as O
    O.lib()
    host = host_cpu()
    n_threads = host["threads"]
    offs = inp["member_offsets"]
    n_members = len(offs) - 1
    if name == "ac_field":
        dt, steps, implicit, interactions = 1e-14, 10000, False, False
    elif name == "equilibrium":
        dt, steps, implicit, interactions = 1e-14, 100000, False, False   # whole runs
    elif name == "implicit_clusters":
        dt, steps, implicit, interactions = 5e-13, 20, True, True
    elif name == "replicas":
        dt, steps, implicit, interactions = 1e-14, 20, False, True
    elif name == "implicit":
        dt, steps, implicit, interactions = 5e-13, 100, True, False
    else:
        dt, steps, implicit, interactions = 1e-14, 200, False, True
    lock = threading.Lock()
    state = {"next": 0, "done": 0}
    t_stop = [0.0]

    def member(i):
        b, e = int(offs[i]), int(offs[i + 1])
        o = O.full_dynamics(inp["radius"][b:e], inp["anisotropy"][b:e], inp["anisotropy_axis"][b:e],
                            inp["magnetisation_direction"][b:e], inp["location"][b:e], 400e3, 0.1,
                            300.0, False, interactions, implicit, 1e-9, dt, steps * dt, 2,
                            int(inp["seeds"][i]), field_shape=inp["field_shape"],
                            field_amplitude=inp["field_amplitude"],
                            field_frequency=inp["field_frequency"], rng="mt")
        return (e - b) * o["stats"].n_steps

    def worker():
        while True:
            with lock:
                if time.perf_counter() > t_stop[0]:
                    return
                i = state["next"] % n_members
                state["next"] += 1
            ps = member(i)
            with lock:
                state["done"] += ps

    def run(nt, secs):
        state["next"], state["done"] = 0, 0
        t0 = time.perf_counter()
        t_stop[0] = t0 + secs
        th = [threading.Thread(target=worker) for _ in range(nt)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        return state["done"] / (time.perf_counter() - t0), state["next"]

    one, m1 = run(1, seconds / 4)
    allc, ma = run(n_threads, seconds)
    what = {"ac_field": "Heun", "implicit": "implicit-midpoint", "clusters": "Heun, dipolar",
            "replicas": "Heun, dipolar (512 particles)", "equilibrium": "Heun",
            "implicit_clusters": "implicit-midpoint, dipolar"}[name]
    return {"value": allc, "unit": "particle-steps/s", "cores": n_threads, "kind": "port",
            "value_1core": one, "host": host,
            "sample": "%d members x %d %s steps each (FP64 oracle in the reference's structure: "
                      "mt19937_64 Wiener draws, dense 3N x 3N diffusion, function-pointer sde) "
                      "on %d threads; 1-core rate from %d members" % (ma, steps + 1, what,
                                                                     n_threads, m1)}


def sar_run(n, precision, save, stop=0, checkpoint="", resume="", noise_tag=0):
    """The full AC-field workload (5 cycles) in one precision; prints progress
    and the final-cycle hysteresis energy through EnsembleResults.

    A run longer than one sitting is split: ``stop`` ends it at that step and
    writes the state to ``checkpoint`` (npz: step, state); ``resume`` starts
    from such a file.  The Philox stream is indexed by step, so the split run
    equals the unsplit one bit for bit."""
    import magpy_amd
    from magpy_amd.results import EnsembleResults
    core = magpy_amd.core
    inp = ac_field_inputs(n)
    f, cycles = 3e5, 5
    print("device:", core.device_info(0), flush=True)
    start = {}
    if resume:
        ck = np.load(resume)
        start = dict(start_step=int(ck["step"]), start_state=ck["state"])
        print(json.dumps({"resume_from_step": start["start_step"]}), flush=True)
    ses = core.Session(precision=precision, trajectories=False, ensemble_sum=True,
                       time_step=1e-14, end_time=cycles / f, max_samples=2000 * cycles + 1,
                       interactions=False, use_implicit=False, noise_tag=noise_tag, **start,
                       **inp)
    total = ses.info()["n_steps"]
    upto = min(stop, total) if stop else total
    t0 = time.perf_counter()
    done = start.get("start_step", 0)
    chunk = 1 << 24
    while done < upto:
        done = ses.advance(min(chunk, upto - done))
        print(json.dumps({"progress_steps": int(done), "of": int(total),
                          "elapsed_s": round(time.perf_counter() - t0, 1)}), flush=True)
    if done < total:
        state, step = ses.state()
        np.savez(checkpoint, step=step, state=state)
        print(json.dumps({"checkpoint": checkpoint, "step": int(step),
                          "wall_s": time.perf_counter() - t0}), flush=True)
        return
    res = ses.finish()
    el = time.perf_counter() - t0
    er = EnsembleResults.from_batch(res, inp["member_offsets"])
    energy = er.final_cycle_energy_dissipated(f)
    if save:
        np.savez(save, time=res["time"], field=res["field"], ensemble_m=res["ensemble_m"],
                 n_members=n)
    print(json.dumps({"sar_run": precision, "noise_tag": noise_tag, "particles": n,
                      "steps": int(res["n_steps"]),
                      "final_cycle_energy_J_per_m3": energy, "sar_W_per_m3": energy * f,
                      "wall_s": el, "kernel_ms": res["kernel_ms"],
                      "nonfinite": int(np.sum(res["status"] != 0))}), flush=True)


# Arithmetic of the sensitivity kernels per particle-step, counted from
# adjoint_kernels.hpp (add/sub/mul = 1, the FP64 gradient sums included):
# one forward Heun step of the fused form is 118 FLOP; the transfer kernel
# recomputes the step (118), its intermediates (75) and runs three adjoint
# channels (3 x 250).
FLOP_ADJ_FORWARD = 118
FLOP_ADJ_TRANSFER = 118 + 75 + 3 * 250


def adjoint_bench(args):
    """Gradient throughput of the differentiation workload: 16,384 particles
    of the AC-field distribution, theta = (K, alpha, H, median r) at
    (23 kJ/m^3, 0.1, 20 mT, 10 nm), checkpoints every 1,000 steps.  A bench
    step is about --block time steps through the recording forward pass (the
    forward kernel plus the adjoint maps of the block's segments; the library
    runs whole blocks of steps, so a step ends on the first block boundary at
    or past its share); the run is warmup + steps such steps long and ends
    with the reverse pass (a scan over the sampling intervals, timed and
    added).  value = particle-steps of the timed steps / (their wall time +
    the reverse pass)."""
    import magpy_amd
    core = magpy_amd.core
    P = args.particles or DEFAULTS["adjoint"][0]
    B = args.block or DEFAULTS["adjoint"][1]
    inp = ac_field_inputs(P)
    shape = inp["radius"] / 10e-9
    ax = inp["anisotropy_axis"]
    theta = np.array([23e3, 0.1, 0.02 / MU0, 10e-9])
    n_blocks = args.warmup + args.steps
    N = B * n_blocks
    ks = np.arange(0, N + 1, max(B // 8, 1), dtype=np.int64)
    print("device:", core.device_info(0), flush=True)
    os.environ["MAGPY_AMD_ADJOINT_STEPS_PER_LAUNCH"] = str(B)
    run = core.Adjoint(shape, ax, ax, inp["seeds"], 400e3, 300.0, 3e5, 1e-14, N, ks, 1000, "fp32")
    run.begin(theta, True)
    done0 = 0
    for b in range(args.warmup):
        done0 = run.forward_advance(B * (b + 1) - done0) if done0 < B * (b + 1) else done0
    info0 = run.info()   # waits for the warmup blocks' maps
    t0 = time.perf_counter()
    done1 = done0
    for b in range(args.steps):
        # returns when the forward kernels are done; the last block's maps
        # overlap the next call's forward kernels
        goal = B * (args.warmup + b + 1)
        if done1 < goal:
            done1 = run.forward_advance(goal - done1)
    info1 = run.info()   # waits for the last block's maps
    t_fwd = time.perf_counter() - t0
    ens, fin, status = run.forward_result()
    w = np.linspace(-1.0, 1.0, len(ks))
    t0 = time.perf_counter()
    grad = run.reverse(w, 1.0)
    t_rev = time.perf_counter() - t0
    info2 = run.info()
    ps = float(P) * (done1 - done0)
    value = ps / (t_fwd + t_rev)
    fwd_kernel_s = (info1["forward_ms"] - info0["forward_ms"]) / 1e3
    rec_kernel_s = (info1["reverse_ms"] - info0["reverse_ms"]) / 1e3
    achieved = FLOP_ADJ_TRANSFER * ps / rec_kernel_s / 1e12
    line = {
        "metric": METRIC, "value": value, "unit": "gradient particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": (t_fwd + t_rev) / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "differentiation: 16,384 AC-field particles (20 mT, 300 kHz, "
                               "dt=1e-14 s), pathwise Heun adjoint, FP32 state, checkpoints "
                               "every 1,000 steps; step = %d time steps of the recording "
                               "forward pass" % B,
                   "particles": P, "time_steps_per_step": B, "run_length_steps": N,
                   "time_steps_timed": int(done1 - done0),
                   "samples": len(ks), "checkpoint_bytes": info2["checkpoint_bytes"],
                   "scratch_bytes": info2["scratch_bytes"], "forward_status": int(status),
                   "gradient_finite": bool(np.all(np.isfinite(grad)))},
        "reverse_pass_ms": t_rev * 1e3,
        "device_s_forward_kernels": fwd_kernel_s, "device_s_adjoint_maps": rec_kernel_s,
        "roofline": {"bound": "valu", "kernel": "path_transfer_kernel<float> (+ folding)",
                     "achieved": achieved, "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS["fp32"], "traffic": None,
                     "flop_per_particle_step": FLOP_ADJ_TRANSFER,
                     "forward_kernel_achieved": FLOP_ADJ_FORWARD * ps / fwd_kernel_s / 1e12},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = adjoint_cpu_baseline(inp, theta)
        line["speedup_vs_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


def replica_accuracy(args):
    """Shared-geometry replicas (one 512-particle cluster, 15% volume
    fraction, 20 mT / 300 kHz, 300 K, 1e5 Heun steps of 1e-14 s) with the
    dipolar field as the BF16x3 product, the FP32 product, FP32 pair sums,
    and FP64 pair sums (the reference), all on the same Philox stream:
    deviation of the ensemble <M_z>(t) (101 samples, relative to the total
    Ms) and rms deviation of the final reduced states from the FP64 run, the partial-cycle energy -mu0 int H d<M_z> of each run
    (1 ns is 3e-4 of a field period, so this is not yet a loop area), and
    the dipolar-field error of each product on the FP64 run's final states.
    --particles sets the replica count (default 256)."""
    import magpy_amd
    core = magpy_amd.core
    R = args.particles or 256
    inp = replica_inputs(R)
    n = 512
    kw = dict(time_step=1e-14, end_time=1e5 * 1e-14, max_samples=101, interactions=True,
              trajectories=False, ensemble_sum=True)
    print("device:", core.device_info(0), flush=True)
    runs = {}
    for name, prec, dip in (("fp64_pairs", "fp64", "pairs"), ("fp32_pairs", "fp32", "pairs"),
                            ("gemm_f32", "fp32", "gemm_f32"),
                            ("gemm_bf16x3", "fp32", "gemm_bf16x3")):
        t0 = time.perf_counter()
        o = core.simulate_ensemble(precision=prec, dipolar=dip, **kw, **inp)
        runs[name] = o
        print(json.dumps({"run": name, "wall_s": time.perf_counter() - t0,
                          "status": int(np.max(o["status"]))}), flush=True)
    ref = runs["fp64_pairs"]
    t = ref["time"]
    h = ref["field"]
    ms_total = 400e3 * n * R

    def energy(o):
        mz = o["ensemble_m"][:, 2] / R         # <M_z> per replica, A/m x particles
        return float(-MU0 * np.trapezoid(h, mz))

    final64 = ref["final_state"].reshape(R, n, 3)     # reduced states
    h64 = core.dipolar_field(inp["location"][:n], inp["radius"][:n], inp["anisotropy"][:n],
                             400e3, final64, mode="pairs")
    scale = np.linalg.norm(h64, axis=(1, 2))
    line = {"replica_accuracy": "%d replicas x 512 particles, 1e5 steps" % R,
            "energy_fp64_J_per_m3_x_particles": energy(ref)}
    for name in ("fp32_pairs", "gemm_f32", "gemm_bf16x3"):
        o = runs[name]
        d_ens = np.max(np.abs(o["ensemble_m"][:, 2] - ref["ensemble_m"][:, 2])) / ms_total
        d_fin = np.sqrt(np.mean(np.sum((o["final_state"] - ref["final_state"]) ** 2, axis=1)))
        entry = {"ens_mz_max_dev_rel": float(d_ens), "final_state_rms_dev": float(d_fin),
                 "energy": energy(o)}
        if name.startswith("gemm"):
            hp = core.dipolar_field(inp["location"][:n], inp["radius"][:n], inp["anisotropy"][:n],
                                    400e3, final64, mode=name)
            entry["field_err_max_rel"] = float(np.max(np.linalg.norm(hp - h64, axis=(1, 2)) / scale))
        line[name] = entry
    line["time_s"] = [float(t[0]), float(t[-1])]
    print(json.dumps(line), flush=True)


def library_gemm(args):
    """The library's GEMM (torch.matmul -> hipBLASLt) at the replica
    product's shape, 4096 x 1536 x 1536: one BF16 pass and FP32, HIP-event
    timing over 50 calls on torch's stream."""
    import torch
    M, N, K = 4096, 1536, 1536
    g = torch.Generator(device="cuda").manual_seed(5)
    a = torch.randn(M, K, device="cuda", generator=g)
    bt = torch.randn(N, K, device="cuda", generator=g).t().contiguous()
    ah, bth = a.bfloat16(), bt.bfloat16()

    def timeit(f, n=50):
        for _ in range(5):
            f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n * 1e3

    out = {"library_gemm": "torch.matmul %dx%dx%d" % (M, N, K)}
    for name, f in (("bf16", lambda: torch.matmul(ah, bth)), ("fp32", lambda: torch.matmul(a, bt))):
        us = timeit(f)
        out[name + "_us"] = us
        out[name + "_tflops"] = 2.0 * M * N * K / us / 1e6
    # this library's BF16x3 product of the replica workload (4,096 states of
    # the 512-particle cluster), 10 calls: each uploads its operands, so only
    # a kernel trace times the product itself (gemm_bf16x3_ksplit_kernel<0>)
    import magpy_amd
    inp = replica_inputs(M)
    n = 512
    rng = np.random.default_rng(3)
    m = rng.normal(size=(M, n, 3))
    m /= np.linalg.norm(m, axis=2, keepdims=True)
    for _ in range(10):
        magpy_amd.core.dipolar_field(inp["location"][:n], inp["radius"][:n], inp["anisotropy"][:n],
                                     400e3, m, mode="gemm_bf16x3")
    out["own_product_calls"] = 10
    print(json.dumps(out), flush=True)


def sar_and_weights(ens, time_s, theta_h, f):
    """SAR = -mu0 f trapz(H, <M_z>) over the last period (magpy/results.py:
    188-217) with H = theta_h sin(2 pi f t); returns (sar, dSAR/d ens,
    direct dSAR/dH) through torch autograd."""
    import torch
    from magpy_amd.adjoint import hysteresis_energy
    e = torch.tensor(ens, dtype=torch.float64, requires_grad=True)
    h = torch.tensor(theta_h, dtype=torch.float64, requires_grad=True)
    t = torch.from_numpy(np.asarray(time_s, dtype=np.float64))
    sar = f * hysteresis_energy(e, h * torch.sin(2 * np.pi * f * t), t, f)
    sar.backward()
    return float(sar.detach()), e.grad.numpy(), float(h.grad)


def gradient_run(run, theta, sched_time, f, precision_tag, chunk=1 << 24, record=True):
    """One forward pass with progress lines; with record, the gradients of
    SAR and final <m_z> (two reverse passes)."""
    t0 = time.perf_counter()
    run.begin(theta, record)
    done = 0
    while done < run.n_steps:
        done = run.forward_advance(chunk)
        print(json.dumps({"pass": precision_tag, "progress_steps": int(done),
                          "of": int(run.n_steps),
                          "elapsed_s": round(time.perf_counter() - t0, 1)}), flush=True)
    ens, fin, status = run.forward_result()
    sar, w, dh = sar_and_weights(ens, sched_time, theta[2], f)
    out = {"sar": sar, "final_mz": fin, "status": int(status),
           "forward_s": time.perf_counter() - t0}
    if record:
        g_sar = run.reverse(w, 0.0)
        out["particle_grad_sar"] = run.particle_gradients()
        g_sar[2] += dh
        out["grad_sar"] = g_sar
        out["grad_final_mz"] = run.reverse(np.zeros(len(w)), 1.0)
        out["total_s"] = time.perf_counter() - t0
    return out


def grad_check(args):
    """Differentiation at scale: (1) the full workload, 16,384 AC-field
    particles over 2 cycles at 300 kHz (6.67e8 steps), FP32 gradients of SAR
    and final <m_z> with respect to (K, alpha, H, median r); (2) the same
    particles under a 30 MHz field (2 cycles = 6.67e6 steps): FP32 and FP64
    gradients against central differences of FP64 runs with common random
    numbers (steps 1e-6 and 5e-7 relative, Richardson-extrapolated).  Over
    this horizon the smallest particles approach thermal switching, where
    the pathwise derivative is large and a 1e-4 step leaves the linear
    regime (the oracle's tangent of one 6.7 nm particle is d m_z/d r =
    -6.97e10; central differences give +8.1e11 at 1e-4, -6.90e10 at 1e-6,
    -6.97e10 at 1e-8), hence the smaller steps."""
    import magpy_amd
    core = magpy_amd.core
    P = args.particles or 16384
    inp = ac_field_inputs(P)
    shape = inp["radius"] / 10e-9
    ax = inp["anisotropy_axis"]
    theta = np.array([23e3, 0.1, 0.02 / MU0, 10e-9])
    print("device:", core.device_info(0), flush=True)
    S = 4001
    os.environ["MAGPY_AMD_ADJOINT_STEPS_PER_LAUNCH"] = str(1 << 20)
    cases = ((3e5, "full"), (3e7, "compressed"))
    if args.grad_check_quick:
        cases = cases[1:]
    elif args.grad_check_full:
        cases = cases[:1]
    for f, tag in cases:
        end = 2.0 / f
        k, n_steps = core.sample_schedule(1e-14, end, S)
        ts = np.arange(S) * (end / (S - 1))
        run32 = core.Adjoint(shape, ax, ax, inp["seeds"], 400e3, 300.0, f, 1e-14, n_steps, k,
                             1000, "fp32")
        r = gradient_run(run32, theta, ts, f, tag + " fp32")
        line = {"grad_check": tag, "field_frequency": f, "particles": P, "steps": int(n_steps),
                "samples": S, "sar_W_per_m3": r["sar"], "final_mz": r["final_mz"],
                "grad_sar": r["grad_sar"].tolist(), "grad_final_mz": r["grad_final_mz"].tolist(),
                "wall_s": r["total_s"], "status": r["status"],
                "gradient_particle_steps_per_s": P * float(n_steps) / r["total_s"],
                "adjoint_info": run32.info()}
        del run32
        if tag == "compressed":
            run64 = core.Adjoint(shape, ax, ax, inp["seeds"], 400e3, 300.0, f, 1e-14, n_steps, k,
                                 1000, "fp64")
            r64 = gradient_run(run64, theta, ts, f, tag + " fp64")
            line["grad_sar_fp64"] = r64["grad_sar"].tolist()
            line["grad_final_mz_fp64"] = r64["grad_final_mz"].tolist()
            fd = {"sar": np.zeros(4), "final_mz": np.zeros(4)}
            for i in range(4):
                vals = {}
                for rel in (1e-6, 5e-7):
                    for sgn in (1, -1):
                        th = theta.copy()
                        th[i] += sgn * rel * theta[i]
                        vals[rel, sgn] = gradient_run(run64, th, ts, f, "fd %d %+g" % (i, sgn * rel),
                                                      record=False)
                for out in ("sar", "final_mz"):
                    c = [(vals[rel, 1][out] - vals[rel, -1][out]) / (2 * rel * theta[i])
                         for rel in (1e-6, 5e-7)]
                    fd[out][i] = (4 * c[1] - c[0]) / 3
            line["fd_sar"] = fd["sar"].tolist()
            line["fd_final_mz"] = fd["final_mz"].tolist()
            line["rel_err_sar"] = (np.abs(r["grad_sar"] - fd["sar"]) / np.abs(fd["sar"])).tolist()
            line["rel_err_final_mz"] = (np.abs(r["grad_final_mz"] - fd["final_mz"]) /
                                        np.abs(fd["final_mz"])).tolist()
            line["rel_err_sar_fp64"] = (np.abs(r64["grad_sar"] - fd["sar"]) /
                                        np.abs(fd["sar"])).tolist()
            # how concentrated the SAR gradient is: the 10 largest particle
            # contributions' share of the total magnitude, and FP32 vs FP64
            # with the top 1 % particles (by either precision) set aside
            p32, p64 = r["particle_grad_sar"], r64["particle_grad_sar"]
            share, rest = [], []
            for i in range(4):
                mag = np.abs(p64[:, i])
                share.append(float(np.sort(mag)[-10:].sum() / mag.sum()))
                k = max(1, P // 100)
                out_ = np.union1d(np.argsort(mag)[-k:], np.argsort(np.abs(p32[:, i]))[-k:])
                keep = np.setdiff1d(np.arange(P), out_)
                a64, a32 = p64[keep, i].sum(), p32[keep, i].sum()
                rest.append(float(abs(a32 - a64) / abs(a64)))
            line["sar_grad_top10_share_fp64"] = share
            line["sar_grad_fp32_vs_fp64_without_top1pct"] = rest
            line["rel_err_final_mz_fp64"] = (np.abs(r64["grad_final_mz"] - fd["final_mz"]) /
                                             np.abs(fd["final_mz"])).tolist()
        print(json.dumps(line), flush=True)


def grad_check_scale(args):
    """Config 4's accuracy at its own scale: 16,384 AC-field particles, 2
    cycles at 300 kHz (6.67e8 steps), 4,002 samples (with 4,001, every third
    sample time is an exact multiple of the step, where the ensemble path's
    schedule, taken in each member's reduced units, can move by one step
    under a perturbation of K or alpha; 4,002 puts no sample on a step).

    phase 'adjoint' (--precision fp64 or fp32; --grad-check-groups G and
    --grad-check-group g run the particles g*P/G .. (g+1)*P/G - 1 only, so
    an FP64 run can be split over sittings): the pathwise gradient of SAR and
    final <m_z> and each particle's part of the SAR gradient, all scaled to
    the whole ensemble (SAR and <m_z> are means over particles, so the
    groups' parts add up), saved to --grad-check-out (npz).

    phase 'fd' (--grad-check-in: the FP64 group files, comma separated;
    --grad-check-in32: the FP32 file): the particles carrying the largest
    parts (the 8 largest per component) and 32 drawn at random are run again
    as independent members at theta +- h theta_i for h in (1e-6, 1e-7,
    1e-8), with their own Wiener streams (common random numbers), in FP64
    through the ensemble path; each particle's SAR contribution
    -mu0 f trapz(H, Ms m_z / P) over the final cycle gives central
    differences.  A particle's converged difference is the one at the
    smaller step of the consecutive pair of steps that agree best (chosen
    from the differences alone); it is compared with the adjoint's part for
    that particle (plus the explicit SAR_i / H for H)."""
    import magpy_amd
    core = magpy_amd.core
    P, f, dt, S = 16384, 3e5, 1e-14, 4002
    if args.grad_check_quick:     # the same check over 2 cycles of a 30 MHz field
        f = 3e7
    end = 2.0 / f
    inp = ac_field_inputs(P)
    shape = inp["radius"] / 10e-9
    ax = inp["anisotropy_axis"]
    theta = np.array([23e3, 0.1, 0.02 / MU0, 10e-9])
    k, n_steps = core.sample_schedule(dt, end, S)
    ts = np.arange(S) * (end / (S - 1))
    print("device:", core.device_info(0), flush=True)
    if args.grad_check_scale == "adjoint":
        prec = args.precision or "fp64"
        G, g = max(1, args.grad_check_groups), args.grad_check_group
        part = np.arange(g * P // G, (g + 1) * P // G)
        scale = len(part) / P
        os.environ["MAGPY_AMD_ADJOINT_STEPS_PER_LAUNCH"] = str(1 << 20)
        run = core.Adjoint(shape[part], ax[part], ax[part], inp["seeds"][part], 400e3, 300.0, f,
                           dt, n_steps, k, 1000, prec)
        r = gradient_run(run, theta, ts, f, "scale %s group %d/%d" % (prec, g, G))
        np.savez(args.grad_check_out or "grad_scale_adjoint_%s_%d.npz" % (prec, g), particles=part,
                 particle_grad_sar=r["particle_grad_sar"] * scale,
                 grad_sar=r["grad_sar"] * scale, grad_final_mz=r["grad_final_mz"] * scale,
                 sar=r["sar"] * scale, final_mz=r["final_mz"] * scale, precision=prec)
        print(json.dumps({"grad_check_scale": "adjoint", "precision": prec, "particles": P,
                          "group": [g, G], "group_particles": int(len(part)),
                          "steps": int(n_steps), "samples": S,
                          "sar_part_W_per_m3": r["sar"] * scale,
                          "grad_sar_part": (r["grad_sar"] * scale).tolist(),
                          "grad_final_mz_part": (r["grad_final_mz"] * scale).tolist(),
                          "wall_s": r["total_s"], "status": r["status"],
                          "adjoint_info": run.info()}), flush=True)
        return
    groups = [np.load(fn) for fn in args.grad_check_in.split(",")]
    pg = np.zeros((P, 4))
    for gr in groups:
        pg[gr["particles"]] = gr["particle_grad_sar"]
    assert sum(len(gr["particles"]) for gr in groups) == P, "FP64 groups must cover every particle"
    grad64 = sum(gr["grad_sar"] for gr in groups)
    gfin64 = sum(gr["grad_final_mz"] for gr in groups)
    sar64 = float(sum(gr["sar"] for gr in groups))
    a32 = np.load(args.grad_check_in32) if args.grad_check_in32 else None
    top = np.unique(np.concatenate([np.argsort(np.abs(pg[:, c]))[-8:] for c in range(4)]))
    rng = np.random.default_rng(7)
    rand = np.sort(rng.choice(np.setdiff1d(np.arange(P), top), size=32, replace=False))
    sel = np.concatenate([top, rand])
    steps = (1e-6, 1e-7, 1e-8)
    variants = [(None, 0.0, 0)] + [(c, h, sg) for c in range(4) for h in steps for sg in (1, -1)]
    M = len(sel) * len(variants)
    radius, kk, alpha, hamp, seeds, axes = [], [], [], [], [], []
    for v in variants:
        th = theta.copy()
        if v[0] is not None:
            th[v[0]] += v[2] * v[1] * theta[v[0]]
        radius.append(shape[sel] * th[3])
        kk.append(np.full(len(sel), th[0]))
        alpha.append(np.full(len(sel), th[1]))
        hamp.append(np.full(len(sel), th[2]))
        seeds.append(inp["seeds"][sel])
        axes.append(ax[sel])
    axes = np.concatenate(axes)
    t0 = time.perf_counter()
    ses = core.Session(member_offsets=np.arange(M + 1, dtype=np.int64),
                       radius=np.concatenate(radius), anisotropy=np.concatenate(kk),
                       anisotropy_axis=axes, magnetisation_direction=axes,
                       location=np.zeros((M, 3)), magnetisation=400e3,
                       damping=np.concatenate(alpha), temperature=300.0, field_shape="sine",
                       field_amplitude=np.concatenate(hamp), field_frequency=f,
                       seeds=np.concatenate(seeds), time_step=dt, end_time=end, max_samples=S,
                       interactions=False, precision="fp64", trajectories=True,
                       ensemble_sum=False)
    done, prev = 0, -1
    while done < n_steps and done != prev:
        prev = done
        done = ses.advance(1 << 24)
        print(json.dumps({"pass": "fd fp64", "progress_steps": int(done), "of": int(n_steps),
                          "elapsed_s": round(time.perf_counter() - t0, 1)}), flush=True)
    out = ses.finish()
    mask = ts >= ts[-1] - 1.0 / f
    npart = len(sel)
    sar_i = np.zeros((len(variants), npart))
    for vi, v in enumerate(variants):
        h_amp = theta[2] * (1 + (v[2] * v[1] if v[0] == 2 else 0.0))
        field = h_amp * np.sin(2 * np.pi * f * ts)
        mz = out["z"][vi * npart:(vi + 1) * npart]          # Ms m_z, A/m
        for j in range(npart):
            sar_i[vi, j] = f * -MU0 * np.trapezoid(field[mask], mz[j][mask] / P)
    fd = np.zeros((len(steps), npart, 4))
    for c in range(4):
        for hi, h in enumerate(steps):
            ip = variants.index((c, h, 1))
            im = variants.index((c, h, -1))
            fd[hi, :, c] = (sar_i[ip] - sar_i[im]) / (2 * h * theta[c])
    # converged difference: the smaller step of the best-agreeing consecutive pair
    pair = np.abs(fd[1:] - fd[:-1]) / np.maximum(np.abs(fd[1:]), 1e-300)   # [2, n, 4]
    pick = np.argmin(pair, axis=0) + 1
    fdc = np.take_along_axis(fd, pick[None], axis=0)[0]
    conv = np.take_along_axis(pair, (pick - 1)[None], axis=0)[0]

    def rel(adj):
        a = adj.copy()
        a[:, 2] += sar_i[0] / theta[2]          # SAR_i is linear in the field amplitude
        return a, np.abs(a - fdc) / np.maximum(np.abs(fdc), 1e-300)

    adj, err = rel(pg[sel])
    nt = len(top)
    scale_abs = np.abs(pg).sum(axis=0)
    line = {"grad_check_scale": "fd", "particles": P, "steps": int(n_steps), "samples": S,
            "field_frequency": f, "checked_particles": int(npart), "top": top.tolist(),
            "random": rand.tolist(), "fd_steps": list(steps), "wall_s": time.perf_counter() - t0,
            "sar_fp64_W_per_m3": sar64, "grad_sar_fp64": grad64.tolist(),
            "grad_final_mz_fp64": gfin64.tolist(),
            "share_of_abs_grad_in_top": (np.abs(pg[top]).sum(axis=0) / scale_abs).tolist(),
            "fd_self_agreement_max_top": conv[:nt].max(axis=0).tolist(),
            "fd_self_agreement_max_random": conv[nt:].max(axis=0).tolist(),
            "max_rel_err_fp64_top": err[:nt].max(axis=0).tolist(),
            "max_rel_err_fp64_random": err[nt:].max(axis=0).tolist(),
            "median_rel_err_fp64_random": np.median(err[nt:], axis=0).tolist(),
            "checked_sum_rel_err_fp64": (np.abs(adj.sum(0) - fdc.sum(0)) /
                                         np.abs(fdc.sum(0))).tolist(),
            "status": int(np.max(out["status"]))}
    # the accuracy criterion: the checked particles' summed gradient within
    # 1e-3 of the summed central differences, every component
    line["sum_within_1e-3_fp64"] = bool(np.all(np.array(line["checked_sum_rel_err_fp64"]) < 1e-3))
    if a32 is not None:
        adj32, err32 = rel(a32["particle_grad_sar"][sel])
        g32 = a32["grad_sar"]
        line["grad_sar_fp32"] = g32.tolist()
        line["grad_sar_fp32_vs_fp64"] = (np.abs(g32 - grad64) / np.abs(grad64)).tolist()
        line["max_rel_err_fp32_top"] = err32[:nt].max(axis=0).tolist()
        line["max_rel_err_fp32_random"] = err32[nt:].max(axis=0).tolist()
        line["median_rel_err_fp32_random"] = np.median(err32[nt:], axis=0).tolist()
        sum32 = np.abs(adj32.sum(0) - fdc.sum(0)) / np.abs(fdc.sum(0))
        line["checked_sum_rel_err_fp32"] = sum32.tolist()
        line["sum_within_1e-3_fp32"] = bool(np.all(sum32 < 1e-3))
    np.savez(args.grad_check_out or "grad_scale_fd.npz",
             sel=sel, fd=fd, adj=adj, sar_i=sar_i)
    print(json.dumps(line), flush=True)


DOM_RUN = dict(time_step=1e-12, end_time=5 / 3e5, max_samples=11, field_shape="sine",
               field_amplitude=0.02 / MU0, field_frequency=3e5)


def dom_inputs(n, seed=2024):
    """the AC-field ensemble's particles (same radii), all in the up state"""
    r = ac_field_inputs(n, seed)["radius"]
    return dict(volume=4.0 / 3.0 * np.pi * r ** 3, anisotropy=np.full(n, 23e3),
                temperature=np.full(n, 300.0), magnetisation=np.full(n, 400e3),
                damping=np.full(n, 0.1), initial_probabilities=np.tile([1.0, 0.0], (n, 1)))


def dom_cpu_baseline(inp, steps_per_particle, seconds=12.0):
    """The oracle's orc_dom_ensemble_dynamics (simulation::dom_ensemble_dynamics
    restated) on the host cores, particle by particle; accepted steps are the
    device run's per-particle counts (the same adaptive algorithm)."""
    from oracle import oracle as O
    O.lib()
    host = host_cpu()
    n_threads = host["threads"]
    P = len(inp["volume"])
    lock = threading.Lock()
    state = {"next": 0, "done": 0}
    t_stop = [0.0]

    def particle(i):
        hk = 2 * inp["anisotropy"][i] / MU0 / inp["magnetisation"][i]
        O.dom_ensemble_dynamics(inp["volume"][i], inp["anisotropy"][i], inp["temperature"][i],
                                inp["magnetisation"][i], inp["damping"][i], "sine",
                                DOM_RUN["field_amplitude"] / hk, DOM_RUN["field_frequency"],
                                inp["initial_probabilities"][i], DOM_RUN["time_step"],
                                DOM_RUN["end_time"], DOM_RUN["max_samples"])
        return int(steps_per_particle[i])

    def worker():
        while True:
            with lock:
                if time.perf_counter() > t_stop[0]:
                    return
                i = state["next"] % P
                state["next"] += 1
            ps = particle(i)
            with lock:
                state["done"] += ps

    def run(nt, secs):
        state["next"], state["done"] = 0, 0
        t0 = time.perf_counter()
        t_stop[0] = t0 + secs
        th = [threading.Thread(target=worker) for _ in range(nt)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        return state["done"] / (time.perf_counter() - t0), state["next"]

    one, m1 = run(1, seconds / 4)
    allc, ma = run(n_threads, seconds)
    return {"value": allc, "unit": "particle-RK45-steps/s", "cores": n_threads, "kind": "port",
            "host": host,
            "value_1core": one,
            "sample": "%d particles' whole runs of the FP64 oracle's dom_ensemble_dynamics "
                      "(function-pointer master equation, reference rk45) on %d threads; "
                      "1-core rate from %d particles" % (ma, n_threads, m1)}


def dom_bench(args):
    import magpy_amd
    core = magpy_amd.core
    n = args.particles or DEFAULTS["dom"][0]
    inp = dom_inputs(n)
    print("device:", core.device_info(0), flush=True)
    for _ in range(args.warmup):
        core.simulate_dom_ensemble(**inp, **DOM_RUN)
    steps, kernel_s, wall, bad = 0, 0.0, 0.0, 0
    per = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        r = core.simulate_dom_ensemble(**inp, **DOM_RUN)
        wall += time.perf_counter() - t0
        steps += int(r["steps"].sum())
        kernel_s += r["kernel_ms"] / 1e3
        bad += int(np.count_nonzero(r["status"]))
        per = r["steps"]
    value = steps / kernel_s
    achieved = FLOP_DOM_PER_STEP * steps / kernel_s / 1e12
    line = {
        "metric": METRIC, "value": value, "unit": "particle-RK45-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": kernel_s / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "discrete orientation model of the 65,536-particle AC-field "
                               "ensemble (two-state master equation, Cash-Karp RK45 FP64, "
                               "tolerance 1e-12, 5 cycles of 20 mT / 300 kHz); step = one whole "
                               "run of all particles",
                   "particles": n, "rk45_steps_per_particle": float(per.mean()),
                   "particles_flagged": bad},
        "pcie_inclusive_value": steps / wall,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["fp64"],
                     "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["fp64"], "traffic": None,
                     "avg_launch_ms": kernel_s / args.steps * 1e3,
                     "flop_per_rk45_step": FLOP_DOM_PER_STEP},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = dom_cpu_baseline(inp, per)
        line["speedup_vs_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


def adjoint_cpu_baseline(inp, theta, seconds=12.0):
    """The oracle's FP64 forward-mode tangent (4 directions) on the host
    cores: gradient particle-steps/s on a bounded sample of the workload."""
    from oracle import oracle as O
    import magpy_amd
    O.lib()
    host = host_cpu()
    n_threads = host["threads"]
    steps, per = 20000, 4
    shape = inp["radius"] / 10e-9
    q, dq = magpy_amd.core.adjoint_reduced(theta, shape, 400e3, 300.0, 1e-14)
    dirs = np.ascontiguousarray(np.transpose(dq, (0, 2, 1)))
    P = len(shape)
    lock = threading.Lock()
    state = {"next": 0, "done": 0}
    t_stop = [0.0]

    def chunk(i):
        b = (i * per) % (P - per + 1)
        sl = slice(b, b + per)
        ax = inp["anisotropy_axis"][sl]
        pr = O.path_problem(ax, ax, q[sl, 0], q[sl, 1], q[sl, 2], q[sl, 3], 3e5 * 1e-14,
                            inp["seeds"][sl], steps, [0, steps])
        O.path_tangent(pr, dirs[sl])
        return per * steps

    def worker():
        while True:
            with lock:
                if time.perf_counter() > t_stop[0]:
                    return
                i = state["next"]
                state["next"] += 1
            ps = chunk(i)
            with lock:
                state["done"] += ps

    def run(nt, secs):
        state["next"], state["done"] = 0, 0
        t0 = time.perf_counter()
        t_stop[0] = t0 + secs
        th = [threading.Thread(target=worker) for _ in range(nt)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        return state["done"] / (time.perf_counter() - t0), state["next"]

    one, m1 = run(1, seconds / 4)
    allc, ma = run(n_threads, seconds)
    return {"value": allc, "unit": "gradient particle-steps/s", "cores": n_threads, "host": host,
            "kind": "port", "value_1core": one,
            "sample": "%d chunks of %d particles x %d steps of the FP64 oracle's forward-mode "
                      "tangent (4 directions: dq/d(K, alpha, H, r)) on %d threads; 1-core rate "
                      "from %d chunks" % (ma, per, steps, n_threads, m1)}


def copy_calibration(nbytes):
    """One device-to-device hipMemcpy of nbytes (a blit kernel on ROCm), the
    known-size reference for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--block", type=int, default=0, help="time steps per bench step")
    ap.add_argument("--particles", type=int, default=0, help="particles (clusters for 'clusters')")
    ap.add_argument("--workload", default="ac_field", choices=sorted(DEFAULTS))
    ap.add_argument("--dipolar", default="auto",
                    choices=["auto", "pairs", "gemm_bf16x3", "gemm_f32"])
    ap.add_argument("--precision", default="")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run one device copy of BYTES (counter calibration)")
    ap.add_argument("--sar-run", default="", choices=["", "fp32", "fp64"],
                    help="run the whole AC-field workload in this precision and report SAR")
    ap.add_argument("--save", default="", help="npz file for the --sar-run time series")
    ap.add_argument("--sar-stop", type=int, default=0,
                    help="end the --sar-run at this step and write --sar-checkpoint")
    ap.add_argument("--sar-checkpoint", default="", help="npz written at --sar-stop")
    ap.add_argument("--sar-resume", default="", help="npz checkpoint to resume --sar-run from")
    ap.add_argument("--grad-check", action="store_true",
                    help="run the full differentiation workload and a finite-difference check")
    ap.add_argument("--replica-accuracy", action="store_true",
                    help="replica runs with each dipolar product against the FP64 pair sums")
    ap.add_argument("--library-gemm", action="store_true",
                    help="time the library GEMM at the replica product's shape")
    ap.add_argument("--grad-check-full", action="store_true",
                    help="with --grad-check: only the full 300 kHz gradient run")
    ap.add_argument("--grad-check-quick", action="store_true",
                    help="with --grad-check: only the finite-difference check")
    ap.add_argument("--grad-check-scale", default="", choices=["", "adjoint", "fd"],
                    help="the differentiation workload's gradient checked at its own scale")
    ap.add_argument("--grad-check-out", default="", help="npz written by --grad-check-scale")
    ap.add_argument("--grad-check-in", default="", help="FP64 adjoint npz for the fd phase")
    ap.add_argument("--grad-check-in32", default="", help="FP32 adjoint npz for the fd phase")
    ap.add_argument("--grad-check-groups", type=int, default=1,
                    help="particle groups of the adjoint phase (one per sitting)")
    ap.add_argument("--grad-check-group", type=int, default=0, help="the group this run takes")
    ap.add_argument("--noise-tag", type=int, default=0,
                    help="Philox stream tag of the --sar-run (another thermal history)")
    args = ap.parse_args()
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
        return
    if args.library_gemm:
        library_gemm(args)
        return
    if args.replica_accuracy:
        replica_accuracy(args)
        return
    if args.grad_check_scale:
        grad_check_scale(args)
        return
    if args.grad_check:
        grad_check(args)
        return
    if args.sar_run:
        sar_run(args.particles or 65536, args.sar_run, args.save, args.sar_stop,
                args.sar_checkpoint, args.sar_resume, args.noise_tag)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build's path runs on a single MI355X "
              "(how the path would shard over several GPUs is a design, not a run)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "adjoint":
        adjoint_bench(args)
        return
    if args.workload == "dom":
        dom_bench(args)
        return
    import magpy_amd
    core = magpy_amd.core
    n = args.particles or DEFAULTS[args.workload][0]
    block = args.block or DEFAULTS[args.workload][1]
    inp, kw, flop_per_step, prec, desc = workload(args.workload, n,
                                                  block * (args.warmup + args.steps) + 1)
    prec = args.precision or prec
    n_particles = int(inp["member_offsets"][-1])
    print("device:", core.device_info(0), flush=True)
    os.environ["MAGPY_AMD_STEPS_PER_LAUNCH"] = str(block)
    ses = core.Session(precision=prec, trajectories=False, ensemble_sum=True,
                       dipolar=args.dipolar, **kw, **inp)
    for _ in range(args.warmup):
        ses.advance(block)
    info0 = ses.info()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ses.advance(block)
    t1 = time.perf_counter()
    info1 = ses.info()
    elapsed = t1 - t0
    value = float(n_particles) * block * args.steps / elapsed
    launches = info1["launches"] - info0["launches"]
    avg_launch_s = (info1["kernel_ms"] - info0["kernel_ms"]) / 1e3 / max(launches, 1)
    steps_per_launch = block * args.steps / max(launches, 1)
    st, step = ses.state()
    line = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": desc + ", step = %d time steps of all particles" % block,
                   "particles": n_particles, "members": len(inp["member_offsets"]) - 1,
                   "time_steps_per_step": block, "run_length_steps": info1["n_steps"],
                   "precision": prec, "state_finite": bool(np.all(np.isfinite(st))),
                   "cluster_block": info1["cluster_block"],
                   "cluster_lanes": info1["cluster_lanes"], "dipolar": info1["dipolar"],
                   "step_index": int(step)},
    }
    if flop_per_step is not None and args.workload == "replicas":
        # two products of (3N)^2 x R per step dominate; BF16 three-term split
        # runs three MFMA products for each
        achieved = flop_per_step * steps_per_launch / avg_launch_s / 1e12
        gemm_mode = info1["dipolar"]
        peak = 2500.0 if gemm_mode == "gemm_bf16x3" else PEAK_TFLOPS["fp32"]
        passes = 3 if gemm_mode == "gemm_bf16x3" else 1
        line["roofline"] = {
            "bound": "mfma" if gemm_mode.startswith("gemm") else "valu", "achieved": achieved,
            "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "mfma_utilisation": passes * achieved / peak, "traffic": None,
            "avg_step_ms": avg_launch_s * 1e3, "flop_per_time_step": flop_per_step}
    elif flop_per_step is not None:
        achieved = flop_per_step * steps_per_launch / avg_launch_s / 1e12
        peak = PEAK_TFLOPS[prec]
        default = (args.workload, n, block, prec) == ("ac_field", 65536, 131072, "fp32")
        line["roofline"] = {
            "bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": TRAFFIC_BYTES_DEFAULT_LAUNCH if default else None,
            "avg_launch_ms": avg_launch_s * 1e3, "launches": launches,
            "flop_per_time_step": flop_per_step}
        if default:
            line["sar_accuracy_recorded"] = SAR_ACCURACY_RECORDED
    else:
        fin = ses.finish()
        # Newton iterations are counted per member (one 3N system per
        # cluster, as the reference counts them); every particle of a
        # member takes part in each of its member's iterations
        n_members = len(inp["member_offsets"]) - 1
        iters = fin["newton_iters"] / (n_members * max(fin["n_steps"], 1))
        flop_step = FLOP_IMPLICIT_PER_STEP + iters * FLOP_IMPLICIT_PER_NEWTON
        if args.workload == "implicit_clusters":
            # one dipolar field evaluation for the guess and one per Newton
            # iterate, 24 FLOP per ordered pair each (the members' mean
            # iteration count applied to every particle)
            sizes = np.diff(inp["member_offsets"])
            pairs_per_particle = float(np.sum(sizes * (sizes - 1))) / n_particles
            flop_step += FLOP_PER_PAIR_STEP / 2 * pairs_per_particle * (1 + iters)
        achieved = flop_step * n_particles * steps_per_launch / avg_launch_s / 1e12
        line["roofline"] = {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS[prec],
                            "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS[prec],
                            "traffic": None, "avg_launch_ms": avg_launch_s * 1e3,
                            "launches": launches, "flop_per_particle_step": flop_step,
                            "flop_per_newton_iteration": FLOP_IMPLICIT_PER_NEWTON}
        line["newton_iters_per_member_step"] = iters
        line["newton_failures"] = fin["newton_failures"]
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.workload, inp)
        line["speedup_vs_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
