"""Throughput of the sLLG Heun ensemble path on one MI355X.

Workload (the 65,536-particle AC-field ensemble): 65,536 independent
particles, radii lognormal (median 10 nm, sigma 0.1), K = 23 kJ/m^3, easy
axes uniform on the sphere, m0 along the easy axis, alpha = 0.1, T = 300 K,
sinusoidal z field 20 mT at 300 kHz, Heun, FP32, dt = 1e-14 s, run length
5 field cycles (1.67e9 steps) with 2,000 samples per cycle reduced on the
device.  Inputs are synthetic and resident in HBM before timing starts.

One bench "step" is one step block: BLOCK time steps of all particles, one
kernel launch.  value = particles * BLOCK * steps / wall time of the timed
blocks (particle-steps/s).  The roofline object divides the algorithmic FP32
work of one launch (278 FLOP per particle-step, counted from the reference's
formulas) by the average launch duration measured with HIP events on the
launch stream.  The CPU baseline times the FP64 oracle (a reference-
structured restatement of simulation::full_dynamics with the reference's
mt19937_64 Wiener source) on a bounded sample of the same workload.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--block B]
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
PEAK_FP32_TFLOPS = 157.3              # MI355X vector FP32 (published)
# HBM bytes of one default launch (65,536 particles x 131,072 steps), from
# rocprofv3 FETCH_SIZE (x2.000) + WRITE_SIZE (x1.000) scaled against a 1 GiB
# device copy; see profiles/heun_indep_fp32_65536_ac_field.md.
TRAFFIC_BYTES_DEFAULT_LAUNCH = 13.30e6


def workload(n, seed=2024):
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


def cpu_baseline(inp, seconds=12.0, steps=10000, chunk=16):
    """Time the oracle on the host cores (threads release the GIL in ctypes)."""
    from oracle import oracle as O
    O.lib()
    n_threads = min(os.cpu_count() or 1, 16)
    dt = 1e-14
    end = steps * dt
    lock = threading.Lock()
    state = {"next": 0, "done": 0}
    t_stop = [0.0]

    def worker():
        while True:
            with lock:
                if time.perf_counter() > t_stop[0]:
                    return
                i0 = state["next"] % len(inp["radius"])
                state["next"] += chunk
            for i in range(i0, min(i0 + chunk, len(inp["radius"]))):
                O.full_dynamics([inp["radius"][i]], [23e3], [inp["anisotropy_axis"][i]],
                                [inp["magnetisation_direction"][i]], [[0, 0, 0]], 400e3, 0.1,
                                300.0, False, True, False, 1e-9, dt, end, 2, int(inp["seeds"][i]),
                                field_shape='sine', field_amplitude=0.02 / MU0,
                                field_frequency=3e5, rng="mt")
                with lock:
                    state["done"] += 1

    def run(nt, secs):
        state["next"], state["done"] = 0, 0
        t0 = time.perf_counter()
        t_stop[0] = t0 + secs
        th = [threading.Thread(target=worker) for _ in range(nt)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        el = time.perf_counter() - t0
        return state["done"] * (steps + 1) / el, state["done"]

    one, n1 = run(1, seconds / 4)
    allc, na = run(n_threads, seconds)
    return {"value": allc, "unit": "particle-steps/s", "cores": n_threads, "kind": "port",
            "value_1core": one,
            "sample": "%d particles x %d Heun steps each (FP64 oracle, mt19937_64 Wiener "
                      "draws, dense 3x3 diffusion, function-pointer sde) on %d threads; "
                      "1-core rate from %d particles" % (na, steps + 1, n_threads, n1)}


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
    ap.add_argument("--block", type=int, default=131072, help="time steps per bench step")
    ap.add_argument("--particles", type=int, default=65536)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run one device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        copy_calibration(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build's path runs on a single MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not run)", file=sys.stderr)
        sys.exit(2)

    import magpy_amd
    core = magpy_amd.core
    n = args.particles
    inp = workload(n)
    cycles, f = 5, 3e5
    end_time = cycles / f
    max_samples = 2000 * cycles + 1
    print("device:", core.device_info(0), flush=True)
    os.environ["MAGPY_AMD_STEPS_PER_LAUNCH"] = str(args.block)
    ses = core.Session(time_step=1e-14, end_time=end_time, max_samples=max_samples,
                       interactions=False, use_implicit=False, precision=args.precision,
                       trajectories=False, ensemble_sum=True, **inp)
    for i in range(args.warmup):
        ses.advance(args.block)
    info0 = ses.info()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ses.advance(args.block)
    t1 = time.perf_counter()
    info1 = ses.info()
    elapsed = t1 - t0
    psteps = float(n) * args.block * args.steps
    value = psteps / elapsed
    launches = info1["launches"] - info0["launches"]
    kms = info1["kernel_ms"] - info0["kernel_ms"]
    avg_launch_s = kms / 1e3 / max(launches, 1)
    flop_per_launch = FLOP_PER_PARTICLE_STEP * float(n) * (args.block * args.steps / max(launches, 1))
    achieved = flop_per_launch / avg_launch_s / 1e12
    st, step = ses.state()
    ok = bool(np.all(np.isfinite(st)))
    line = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": "65,536-particle AC-field Heun ensemble (20 mT, 300 kHz, "
                               "dt=1e-14 s), step = %d time steps of all particles" % args.block,
                   "particles": n, "time_steps_per_step": args.block,
                   "run_length_steps": info1["n_steps"], "samples": max_samples,
                   "precision": args.precision, "state_finite": ok, "step_index": int(step)},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": (TRAFFIC_BYTES_DEFAULT_LAUNCH
                                 if (n, args.block, args.precision) == (65536, 131072, "fp32")
                                 else None),
                     "kernel": "heun_indep_kernel<float, sine>",
                     "avg_launch_ms": avg_launch_s * 1e3, "launches": launches,
                     "flop_per_particle_step": FLOP_PER_PARTICLE_STEP},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(inp)
        line["speedup_vs_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
