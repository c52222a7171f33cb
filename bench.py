"""Coupled EHD step benchmark on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on):
256 x 256 x 128 box, FP32-shifted populations, T=400, C=10, M=10, Fe=4000,
Pr=10, TRT Lambda=1/4, hydrostatic start plus seeded 1e-3 velocity noise.
A "step" is one full coupled step: fused LBM collide/stream/boundary for f
and h, then the Poisson solve (x,y FFTs, per-mode tridiagonal z solve,
inverse FFTs).  MLUPS = nx*ny*nz*steps / (1e6 * seconds), walls included
(main.cu:256-258).

Prints one JSON line.  roofline: the fused LBM kernel (the dominant kernel),
algorithmic bytes 440 B per node (2 x 27 populations read and written in
FP32, q written, phi read) over its average launch time, taken from HIP
events recorded on the box's stream around the launch of every 8th step of
the timed region (--profile-every).  copy_ceiling_gbs: a 2 GiB float4
device-to-device copy in the same process (ehd_copy_ceiling), plain and
nontemporal, the box's practical HBM ceiling.  fp64: the same box in the
reference's precision (FP64 populations, 880 B per node for the LBM kernel).
cpu_baseline: the FP64 CPU oracle (OpenMP, one thread per physical core,
bound) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "ehd-convection-3d_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
LBM_BYTES_PER_NODE = {1: 440.0, 0: 880.0}     # FP32S, FP64
STEP_BYTES_PER_NODE = {1: 448.0, 0: 896.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--precision", choices=("fp32s", "fp64"), default="fp32s")
    ap.add_argument("--T", type=float, default=400.0, help="electric Rayleigh number of the box")
    ap.add_argument("--seed", type=int, default=2, help="seed of the 1e-3 velocity noise")
    ap.add_argument("--cpu-steps", type=int, default=8,
                    help="oracle steps per timed run of the CPU sample, three runs (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle OpenMP threads (0: OMP_NUM_THREADS, else the cores this process may use)")
    ap.add_argument("--no-profile", action="store_true", help="no per-launch events (value only)")
    ap.add_argument("--profile-every", type=int, default=8,
                    help="HIP events around the LBM and Poisson launches of every n-th step (each event record idles the stream for a few microseconds; every launch does the same work)")
    ap.add_argument("--fp64-steps", type=int, default=40,
                    help="steps of the same box in FP64 for the fp64 sub-object (0: skip)")
    ap.add_argument("--workload", choices=("box", "ensemble", "fno-train"), default="box",
                    help="box: one 256x256x128 box (the headline line); ensemble: the 48-box sweep in one batch; "
                         "fno-train: training steps of the warm-start surrogate (64^3, width 32, 12 modes, batch 8)")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join("profiles", "lbm_traffic.json")


def traffic_from_profiles(workload: str):
    """HBM bytes per LBM launch of this workload as two rocprofv3 PMC runs of
    the same command measured them (FETCH_SIZE, then WRITE_SIZE, each scaled
    by what the counter reads for a copy of known size), committed in
    profiles/lbm_traffic.json with the command that made them; None if the
    file holds no entry for the workload."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            data = json.load(f)
        return data.get(workload, {}).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def physical_cores(cpus):
    """Distinct physical cores among the logical CPUs `cpus` (sysfs topology)."""
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
        except OSError:
            return len(cpus)
        cores.add((pkg, core))
    return len(cores)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, workload):
    # one OpenMP thread per physical core this process may use, bound to its
    # core (the OpenMP runtime reads these when the oracle library loads)
    cpus = sorted(os.sched_getaffinity(0))
    cores = physical_cores(cpus)
    limit = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = args.cpu_threads or (min(cores, limit) if limit else cores)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    from oracle import pyoracle as O

    O.set_threads(threads)
    p = O.make_params(args.nx, args.ny, args.nz, T=args.T, M=10.0, C_=10.0, Fe=4000.0, lam=0.25)
    b = O.Box(p)
    b.init_hydrostatic(args.seed, 1e-3)
    b.init_equilibrium()
    b.step(1)                       # warm-up step
    n = args.cpu_steps
    S = args.nx * args.ny * args.nz
    rates, total = [], 0.0
    for _ in range(3):              # median of three timed runs
        t0 = time.perf_counter()
        b.step(n)
        dt = time.perf_counter() - t0
        total += dt
        rates.append(S * n / dt / 1e6)
    return {"value": round(sorted(rates)[1], 4), "unit": "MLUPS", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "logical_cpus": len(cpus), "physical_cores": cores,
            "binding": "OMP_PROC_BIND=close OMP_PLACES=cores",
            "sample": f"FP64 oracle, {workload}, median of 3 runs of {n} coupled steps after 1 warm-up step, "
                      f"{total:.1f} s in all"}


def ensemble_params(E):
    """48 boxes of a lateral-wavelength / T / C sweep: nx = ny in {48..128},
    nz in {32, 64}, T log-spaced 150-600, C in {5, 10, 20}, M = 10, FP32S."""
    out = []
    for b in range(48):
        n = (48, 64, 80, 96, 128)[b % 5]
        nz = (32, 64)[(b // 5) % 2]
        T = 150.0 * 4.0 ** (b / 47.0)
        out.append(E.make_params(n, n, nz, T=T, C_=(5.0, 10.0, 20.0)[b % 3], M=10.0, Fe=4000.0, Pr=10.0, lam=0.25,
                                 precision=E.FP32S, init_mode=E.INIT_HYDROSTATIC, seed=100 + b, amp=1e-3))
    return out


def run_ensemble(args):
    import ehd_amd as E

    params = ensemble_params(E)
    nodes = sum(p.nx * p.ny * p.nz for p in params)
    with E.Batch(params, flags=E.BATCH_GRAPH | E.BATCH_TUNE) as bt:
        bt.init()
        bt.step(args.warmup)
        t0 = time.perf_counter()
        bt.step(args.steps)
        elapsed = time.perf_counter() - t0
        tu = bt.tuning()
    # the same boxes stepped one after another (one box per call, no graph)
    nseq = max(1, args.steps // 10)
    boxes = [E.Box(p) for p in params]
    for b in boxes:
        b.init()
        b.step(2)
    for b in boxes:
        b.sync()
    t0 = time.perf_counter()
    for b in boxes:
        b.step(nseq)
    for b in boxes:
        b.sync()
    seq = time.perf_counter() - t0
    for b in boxes:
        b.close()
    line = {
        "metric": "ensemble coupled-step MLUPS", "value": round(nodes * args.steps / elapsed / 1e6, 2), "unit": "MLUPS",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"48-box ensemble, {nodes} nodes, FP32S, one batch (graph + autotuned shapes)"},
        "one_box_at_a_time_MLUPS": round(nodes * nseq / seq / 1e6, 2),
        "tuning": {"lbm_grid": tu.lbm_grid, "lbm_tiles": tu.lbm_tiles, "chains": tu.chains,
                   "step_ms_one_launch_vs_chains": [round(tu.step_ms[0], 4), round(tu.step_ms[1], 4)],
                   "groups": [[tu.group[i].nx, tu.group[i].nboxes, tu.group[i].x_lines, tu.group[i].y_lines]
                              for i in range(tu.ngroups)]},
    }
    print(json.dumps(line), flush=True)


def run_fno_train(args):
    """Adam steps of the surrogate FNO (64^3 grid, width 32, 12 modes per
    axis, 4 layers, batch 8) on a batch resident in HBM.  FLOPs and bytes are
    the GEMMs' algorithmic work as the library counts it per step."""
    import numpy as np
    from ehd_amd.fno import FNO

    B, n = 8, 64
    rng = np.random.default_rng(0)
    X = rng.standard_normal((B, 6, n, n, n)).astype(np.float32)
    Y = rng.standard_normal((B, 5, n, n, n)).astype(np.float32)
    net = FNO(n, n, n, width=32, modes=(12, 12, 12), layers=4, max_batch=B)
    net.set_params(net.init_params(seed=1))
    net.train_step(X, Y, 1e-4)             # loads the batch into HBM
    for _ in range(max(args.warmup, 1)):
        net.train_step(None, None, 1e-4)
    net.gemm_work(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.train_step(None, None, 1e-4)
    elapsed = time.perf_counter() - t0
    w = net.gemm_work(reset=True)
    ms = elapsed * 1e3 / args.steps
    line = {
        "metric": "surrogate training step", "value": round(ms, 3), "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 products, f32 accumulate", "data": "synthetic",
        "config": {"workload": f"FNO Adam step, {n}^3 grid, width 32, modes 12x12x12, 4 layers, batch {B}, "
                               "batch resident in HBM"},
        "gemm": {"gflop_per_step": round(w.flops / args.steps / 1e9, 3),
                 "gbytes_per_step": round(w.bytes / args.steps / 1e9, 3),
                 "launches_per_step": w.launches // args.steps,
                 "tflops_over_whole_step": round(w.flops / elapsed / 1e12, 2),
                 "peak_bf16_tflops": 2500.0},
    }
    net.close()
    print(json.dumps(line), flush=True)


def run_box(E, args, prec, steps, warmup, before_warmup=None):
    """One box of the workload stepped warmup + steps times; the timed region
    is the `steps` coupled steps (inputs resident in HBM), bracketed by stream
    synchronisation.  before_warmup (optional) runs after the box's set-up,
    before the warm-up steps.  Returns wall time, the sampled kernel times and
    the diagnostics."""
    nx, ny, nz = args.nx, args.ny, args.nz
    params = E.make_params(nx, ny, nz, T=args.T, C_=10.0, M=10.0, Fe=4000.0, Pr=10.0, lam=0.25,
                           precision=prec, init_mode=E.INIT_HYDROSTATIC, seed=args.seed, amp=1e-3)
    with E.Box(params) as box:
        box.init()
        box.sync()
        if before_warmup is not None:
            before_warmup()
        box.step(warmup)
        box.sync()
        if not args.no_profile:
            box.profile_enable(True, every=args.profile_every)
        box.profile_reset()
        t0 = time.perf_counter()
        box.step(steps)
        box.sync()
        elapsed = time.perf_counter() - t0
        prof = box.profile_read()
        d = box.diagnostics()
    return elapsed, prof, d


def roofline(E, args, prec, prof, workload):
    lbm_ms = prof.lbm_ms / prof.lbm_launches
    S = args.nx * args.ny * args.nz
    achieved = LBM_BYTES_PER_NODE[prec] * S / (lbm_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(workload),
            "traffic_source": TRAFFIC_FILE,
            "kernel": "lbm_f32s_kernel" if prec == E.FP32S else "lbm_f64_kernel",
            "bytes_per_node": LBM_BYTES_PER_NODE[prec], "avg_launch_ms": round(lbm_ms, 4)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the coupled step, the ensemble "
              "and the surrogate run on a single GPU", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        run_ensemble(args)
        return
    if args.workload == "fno-train":
        run_fno_train(args)
        return
    import ehd_amd as E

    prec = E.FP32S if args.precision == "fp32s" else E.FP64
    nx, ny, nz = args.nx, args.ny, args.nz
    S = nx * ny * nz
    workload = f"coupled EHD step {nx}x{ny}x{nz} {args.precision}"
    # the box's own practical ceiling, a 2 GiB float4 copy (plain and
    # nontemporal) in the same process, runs between the box's set-up and its
    # warm-up steps, so that the GPU comes to the warm-up busy, not idle
    ceiling = []
    elapsed, prof, d = run_box(E, args, prec, args.steps, args.warmup,
                               before_warmup=lambda: ceiling.extend(E.copy_ceiling(1 << 31, 20)))
    plain, nt = ceiling
    mlups = S * args.steps / elapsed / 1e6
    ms_step = elapsed * 1e3 / args.steps
    line = {
        "metric": "coupled-step MLUPS", "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == E.FP32S else "f64", "data": "synthetic",
        "config": {"workload": workload, "T": args.T, "C": 10, "M": 10, "Fe": 4000, "Pr": 10,
                   "lambda": 0.25, "init": f"hydrostatic + 1e-3 noise, seed {args.seed}"},
        "hbm_frac_step": round(STEP_BYTES_PER_NODE[prec] * S * args.steps / elapsed / 1e9 / HBM_PEAK_GBS, 4),
    }
    if not args.no_profile and prof.lbm_launches > 0:
        line["roofline"] = roofline(E, args, prec, prof, workload)
        line["kernel_ms"] = {"lbm": line["roofline"]["avg_launch_ms"],
                             "poisson": round(prof.poisson_ms / max(prof.poisson_calls, 1), 4)}
    line["diag"] = {"current": d.current, "umax": d.umax}
    line["copy_ceiling_gbs"] = {"plain": round(plain, 1), "nontemporal": round(nt, 1)}
    if "roofline" in line:
        line["roofline"]["frac_of_copy_ceiling"] = round(line["roofline"]["achieved"] / max(plain, nt), 4)
    if prec == E.FP32S and args.fp64_steps > 0:
        # the reference's own precision on the same workload
        e64, p64, _ = run_box(E, args, E.FP64, args.fp64_steps, 5)
        sub = {"ms_per_step": round(e64 * 1e3 / args.fp64_steps, 4),
               "MLUPS": round(S * args.fp64_steps / e64 / 1e6, 2), "steps": args.fp64_steps,
               "hbm_frac_step": round(STEP_BYTES_PER_NODE[E.FP64] * S * args.fp64_steps / e64 / 1e9
                                      / HBM_PEAK_GBS, 4)}
        if not args.no_profile and p64.lbm_launches > 0:
            r = roofline(E, args, E.FP64, p64, f"coupled EHD step {nx}x{ny}x{nz} fp64")
            sub["roofline"] = {k: r[k] for k in ("kernel", "achieved", "frac", "avg_launch_ms", "traffic",
                                                 "bytes_per_node")}
        line["fp64"] = sub
    if args.cpu_steps > 0:
        line["cpu_baseline"] = cpu_baseline(args, workload)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
