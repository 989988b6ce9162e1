"""Benchmark of the MD step on one MI355X.

Workload (default): 2,048,000-atom LJ fluid, 80^3 FCC cells, rho*=0.8442,
T*=1.0, rc=2.5, skin 0.3, Langevin gamma=1.0, dt=0.005, FP32 pair
arithmetic with 64-bit fixed-point accumulation, FP64 state.  The same
workload is also timed with the plain FP64 kernels; the FP32 forces are
compared with the FP64 ones at one configuration.  A bounded sample of the
same fluid runs on the host CPU (the oracle's clean-room FP64 restatement of
the reference step, one thread) for the cpu_baseline entry.

Prints one JSON line.  Only --gpus 1 is supported (one GPU per box).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "caviar-v2_amd"))

from cv2md import device as D  # noqa: E402
from cv2md.systems import lj_fluid  # noqa: E402

METRIC = "atom-timesteps/s, 2M-atom LJ: FP32+fixed-point vs FP64 GPU at force err ≤1e-6"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(ncell):
    return lj_fluid(ncell, rho=0.8442, temperature=1.0, seed=1003, integrator=1, gamma=1.0)


FORCE_SAMPLE = 8   # HIP events around every 8th force launch of the timed steps
# Before the warm-up steps, force passes at the initial positions (no change
# of state: the same accelerations each time) keep the GPU busy for this
# long, so that its clocks have left the idle state of the host-side setup
# before anything is timed; both forms get the same.
CLOCK_WARMUP_S = 0.2
PHASE_STEPS = 200  # instrumented steps (events around every kernel group) after the timed run


def run_gpu(s, precision, steps, warmup, seed=7):
    """Timed Langevin run; the force launches of the timed steps are sampled
    with HIP events (every FORCE_SAMPLE-th, on the context's stream), and a
    short run with events around every kernel group follows for the phase
    breakdown (events cost a few microseconds each, so that run is not the
    timed one)."""
    n = len(s["pos"])
    with D.Context.from_sysdef(s, precision=precision) as c:
        c.compute_forces()
        t_w = time.perf_counter()
        while time.perf_counter() - t_w < CLOCK_WARMUP_S:
            c.compute_accelerations()
        c.run_langevin(warmup, s["dt"], s["gamma"], s["kbt"], seed)
        c.synchronize()
        c.set_timing(True, period=FORCE_SAMPLE)
        st0 = c.stats()
        t0 = time.perf_counter()
        c.run_langevin(steps, s["dt"], s["gamma"], s["kbt"], seed)
        c.synchronize()
        t1 = time.perf_counter()
        tm = c.timing()
        st = c.stats()
        th = c.thermo(n_r_df=0)
        c.set_timing(True)
        c.run_langevin(PHASE_STEPS, s["dt"], s["gamma"], s["kbt"], seed + 1)
        ph = c.timing()
        c.set_timing(False)
        pos = c.download()[0]
    el = t1 - t0
    return dict(value=n * steps / el, ms_per_step=1e3 * el / steps, timing=tm,
                phases={k: ph[k] / PHASE_STEPS for k in ("force_ms", "build_ms", "integrate_ms")},
                rebuilds=st["rebuilds"] - st0["rebuilds"], mean_list=st["list_entries"] / n,
                max_list=st["max_neighbors"], capacity=st["capacity"],
                temperature=2 * th["ke"] / (3 * n), pos=pos)


def force_error(s, pos):
    """Accelerations of the timed FP32+fixed-point kernel (cv2_compute_accelerations:
    the branchless task kernel the step loop runs) against the FP64 kernel
    at one configuration (the FP64 kernel agrees with the FP64 oracle to
    1e-12, tests/test_gpu_parity.py)."""
    s2 = dict(s, pos=pos)
    with D.Context.from_sysdef(s2, precision=D.CV2_PREC_FP64) as c:
        c.compute_forces()
        ref = c.download()[2]
    with D.Context.from_sysdef(s2, precision=D.CV2_PREC_FP32_FIXED) as c:
        c.compute_accelerations()
        a32 = c.download()[2]
    d = np.linalg.norm(a32 - ref, axis=1)
    m = np.linalg.norm(ref, axis=1)
    rms = float(np.sqrt((m ** 2).mean()))
    return dict(max_rel=float((d / np.maximum(m, rms)).max()), rms_rel=float(np.sqrt((d ** 2).mean()) / rms),
                metric="|a32-a64|/max(|a64|, rms|a64|), per atom; a32 from the timed kernel")


CPU_NCELL = 40      # CPU sample: 256,000 atoms of the same fluid (40^3 FCC cells)
CPU_MIN_STEPS = 20


def cpu_baseline(budget_s):
    """Bounded sample of the same workload on one host core: the oracle's
    clean-room FP64 restatement of the reference step (oracle/cv2_oracle.c,
    bit-identical to the compiled reference on the committed fixtures), on
    a 256,000-atom box of the same fluid, at least CPU_MIN_STEPS consecutive
    Langevin steps including at least one list rebuild."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    s = dict(workload(CPU_NCELL), use_cell_list=1)
    n = len(s["pos"])
    rng = np.random.default_rng(5)
    eta = [rng.normal(size=(n, 3)) for _ in range(4)]  # noise drawn outside the timed loop
    with O.OracleSim(s) as sim:
        t_setup = time.perf_counter()
        sim.setup()
        t_setup = time.perf_counter() - t_setup
        steps = rebuilds = 0
        el = 0.0
        while True:
            e = eta[steps % len(eta)]
            t0 = time.perf_counter()
            rebuilds += int(sim.step(e))
            el += time.perf_counter() - t0
            steps += 1
            if steps >= CPU_MIN_STEPS and rebuilds >= 1 and (el > budget_s or steps >= 60):
                break
            if steps >= 200:
                break
    return dict(value=n * steps / el, unit="atom-timesteps/s", cores=1, kind="port",
                sample=f"{steps} consecutive Langevin steps ({rebuilds} list rebuilds) of a {n}-atom box of "
                       f"the same fluid ({CPU_NCELL}^3 FCC cells, rho*=0.8442, T*=1.0, seed 1003) after an "
                       f"untimed setup ({t_setup:.1f} s); the oracle's FP64 restatement of the reference "
                       f"step (cell-list Verlet path, rebuild on wrap and ghost changes as the reference "
                       f"does), 1 thread, 1 process (MPI ranks: 1)",
                steps=steps, rebuilds=rebuilds, seconds=el,
                cpu=platform.processor() or platform.machine(), nproc=os.cpu_count())


def run_ensemble(steps, seed=1004):
    """Batched ensemble: 512 independent LJ systems (N log-uniform in
    [500, 20000], rho in [0.60, 0.90], T* in [0.8, 2.0]; rc 2.5, skin 0.3),
    FP32+fixed-point, velocity Verlet dt=0.005; launch shapes and rebuild
    policy from the autotuner; each step one captured graph."""
    from cv2md.systems import ensemble_systems

    systems = ensemble_systems(512, seed=seed)
    natoms = sum(len(x["pos"]) for x in systems)
    with D.Ensemble.from_systems(systems, rc=2.5, nl_cutoff=2.5, nl_coef=1.12) as e:
        t0 = time.perf_counter()
        tuned = e.tune(0.005, trial_steps=20)
        t_tune = time.perf_counter() - t0
        e.run_vv(20, 0.005)  # warm-up after the tuner's restore
        st0 = e.stats()
        t0 = time.perf_counter()
        e.run_vv(steps, 0.005)
        el = time.perf_counter() - t0
        st = e.stats()
    return dict(value=natoms * steps / el, unit="atom-timesteps/s", ms_per_step=1e3 * el / steps,
                steps=steps, systems=len(systems), natoms=natoms,
                natoms_min=min(len(x["pos"]) for x in systems),
                natoms_max=max(len(x["pos"]) for x in systems),
                tuning=tuned, tune_seconds=t_tune,
                system_builds=st["system_builds"] - st0["system_builds"],
                build_steps=st["build_steps"] - st0["build_steps"],
                mean_list_length=st["list_entries"] / natoms, graph_nodes=st["graph_nodes"],
                config="512 LJ systems, N log-uniform in [500, 20000], rho* U[0.60, 0.90], T* U[0.8, 2.0], "
                       "random/jittered FCC, rc=2.5, skin 0.3, FP32 + 64-bit fixed point, velocity Verlet "
                       "dt=0.005, one captured HIP graph per step")


def run_share(steps, warmup, seed=11):
    """One MPI domain's share with its halo refreshed from the host every
    step: 1,000,000-atom Kob-Andersen mixture (rho* = 1.2, 80:20,
    rc_ij = 2.5 sigma_ij, 2.8 sigma halo) from random positions, FP64
    jump_fix relaxation (untimed), then Langevin quench steps (T* = 0.5,
    gamma = 0.5, dt = 0.003), FP32 + fixed point.  A timed step includes the
    device halo packing, the device -> host -> device copies of the packets
    and the host's forwarding (PCIe-inclusive)."""
    from cv2md.share import Share, ka_random

    s = ka_random(1000000)
    n = len(s["pos"])
    with Share(s, precision=D.CV2_PREC_FP32_FIXED) as sh:
        t0 = time.perf_counter()
        sh.relax()
        t_relax = time.perf_counter() - t0
        sh.run(warmup, 0.003, True, 0.5, 0.5, seed)
        sh.c.synchronize()
        st0 = sh.c.stats()
        sh.exchange_seconds, sh.pack_seconds, sh.unpack_seconds, sh.exchanges = 0.0, 0.0, 0.0, 0
        t0 = time.perf_counter()
        sh.run(steps, 0.003, True, 0.5, 0.5, seed)
        sh.c.synchronize()
        el = time.perf_counter() - t0
        st = sh.c.stats()
        th = sh.c.thermo(0)
        ex_ms = 1e3 * sh.exchange_seconds / max(sh.exchanges, 1)
        pack_ms = 1e3 * sh.pack_seconds / max(sh.exchanges, 1)
        unpack_ms = 1e3 * sh.unpack_seconds / max(sh.exchanges, 1)
        ghosts = sh.halo_atoms
    return dict(value=n * steps / el, unit="atom-timesteps/s", ms_per_step=1e3 * el / steps,
                exchange_ms_per_step=ex_ms, pack_ms_per_step=pack_ms, unpack_ms_per_step=unpack_ms,
                steps=steps, warmup=warmup, natoms=n, ghosts=ghosts,
                rebuilds=st["rebuilds"] - st0["rebuilds"], build_kind=st["build_kind"],
                relax_seconds=t_relax, temperature_end=2.0 * th["ke"] / (3.0 * n),
                config="Kob-Andersen 80:20, 1,000,000 owned atoms, rho*=1.2, rc_ij=2.5 sigma_ij, 2.8 sigma "
                       "halo packed on the device and forwarded by the host every step (1x1x1 periodic "
                       "process grid); random start, FP64 jump_fix relaxation, Langevin quench T*=0.5, "
                       "gamma=0.5, dt=0.003; FP32 + 64-bit fixed point; PCIe-inclusive")


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_lj_f32_task_lj2m_traffic.txt")


def committed_traffic():
    """HBM bytes per launch of the force kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel scaled),
    or None when there is none."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


COUNTERS_FILE = os.path.join(ROOT, "profiles", "k_lj_f32_task_lj2m_counters.txt")


def committed_counters():
    """VALU busy and L2 hit rate of the force kernel from the committed
    rocprofv3 --pmc summary, or Nones."""
    out = {"valu_busy": None, "l2_hit": None}
    try:
        with open(COUNTERS_FILE) as f:
            for line in f:
                k = line.split(":", 1)[0]
                if k in out:
                    out[k] = float(line.split(":", 1)[1])
    except OSError:
        pass
    return out


def traffic_probe(a):
    """Workload for the rocprofv3 --pmc traffic runs: a few force launches on
    the melted 2M-atom configuration, then a plain copy of known size (the
    calibration kernel: reads and writes COPY_BYTES each)."""
    import torch

    s = workload(a.ncell)
    with D.Context.from_sysdef(s, precision=D.CV2_PREC_FP32_FIXED) as c:
        c.compute_forces()
        c.run_langevin(50, s["dt"], s["gamma"], s["kbt"], 7)
        for _ in range(5):
            c.compute_forces()
        c.synchronize()
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 2.0, out=dst)
    torch.cuda.synchronize()
    log(f"traffic probe done: {len(s['pos'])} atoms, copy {COPY_BYTES} B")


COPY_BYTES = 256 << 20


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the workload's own length: 5000 Langevin steps from the FCC start (the
    # first ~600 melt the lattice and run slower on both forms' kernels)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--ncell", type=int, default=80, help="FCC cells per edge (80 -> 2,048,000 atoms)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ens-steps", type=int, default=2000)
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the batched-ensemble workload and print its JSON (for profiling)")
    ap.add_argument("--no-share", action="store_true")
    ap.add_argument("--share-steps", type=int, default=100)
    ap.add_argument("--share-only", action="store_true",
                    help="run only the host-halo share workload and print its JSON (for profiling)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run only the force kernel and a calibration copy (for rocprofv3 --pmc)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}): one GPU holds the whole box, "
              "and no sharding across GPUs is built", file=sys.stderr)
        sys.exit(2)

    if a.traffic_probe:
        traffic_probe(a)
        return
    if a.share_only:
        print(json.dumps(run_share(a.share_steps, 10)), flush=True)
        return
    if a.ensemble_only:
        r = run_ensemble(a.ens_steps)
        print(json.dumps({k: r[k] for k in ("ms_per_step", "value", "tuning")}), flush=True)
        return
    s = workload(a.ncell)
    n = len(s["pos"])
    log(f"workload: {n} atoms, L={s['hi'][0]:.4f}")
    r32 = run_gpu(s, D.CV2_PREC_FP32_FIXED, a.steps, a.warmup)
    log(f"fp32+fx: {r32['ms_per_step']:.4f} ms/step, rebuilds {r32['rebuilds']}, <n> {r32['mean_list']:.2f}")
    r64 = None
    if not a.no_fp64:
        r64 = run_gpu(s, D.CV2_PREC_FP64, a.steps, a.warmup)
        log(f"fp64: {r64['ms_per_step']:.4f} ms/step, rebuilds {r64['rebuilds']}")
    ferr = force_error(s, r32["pos"])
    log(f"force error: {ferr}")

    tm = r32["timing"]
    t_force = tm["force_ms"] / max(tm["force_calls"], 1) * 1e-3
    nl = r32["mean_list"]
    # The timed force launch is k_lj_f32_task with the integrator fused into
    # its epilogue.  Per atom: q (int4) + nnbr + 16-bit list slots, then the
    # integrator: pos (double4) read + write, vel (3 x f64) read + write,
    # pos_ref read, id read, next q written; the staged neighbourhoods
    # re-read the same q array (counted once), and the task tables are 292
    # bytes per task of about 224 atoms.
    b_force = n * (16 + 4 + 2 * nl + 32 + 32 + 24 + 24 + 32 + 4 + 16) + 292 * (n / 224.0)
    achieved = b_force / t_force / 1e9
    out = {
        "metric": METRIC,
        "value": r32["value"],
        "unit": "atom-timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"LJ fluid {n} atoms ({a.ncell}^3 FCC), rho*=0.8442, T*=1.0, rc=2.5, skin 0.3, "
                        "Langevin gamma=1.0, dt=0.005; FP32 pair arithmetic, 64-bit fixed-point "
                        "accumulation, FP64 state",
            "natoms": n,
            "precision": "fp32+fixed64",
            "mean_list_length": nl,
            "list": "full (both directions), 16-bit task-slot ELL (8-entry groups); cell-row "
                    "neighbourhoods staged in LDS",
            "rebuilds_in_timed_steps": r32["rebuilds"],
            "clock_warmup_s": CLOCK_WARMUP_S,
        },
        "fp64_gpu": None if r64 is None else {"value": r64["value"], "ms_per_step": r64["ms_per_step"],
                                             "force_ms": r64["timing"]["force_ms"] / max(r64["timing"]["force_calls"], 1),
                                             "rebuilds_in_timed_steps": r64["rebuilds"]},
        "speedup_fp32fx_vs_fp64": None if r64 is None else r32["value"] / r64["value"],
        "force_error_vs_fp64": ferr,
        "phase_ms_per_step": r32["phases"],
        "temperature_end": r32["temperature"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": committed_traffic(),
                     "kernel": "k_lj_f32_task (integrator fused)", "bytes_per_launch": b_force,
                     "avg_launch_ms": t_force * 1e3, **committed_counters()},
        "cpu_baseline": None,
        "ensemble": None,
        "share": None,
    }
    if not a.no_ensemble:
        out["ensemble"] = run_ensemble(a.ens_steps)
        log(f"ensemble: {out['ensemble']['ms_per_step']:.4f} ms/step, {out['ensemble']['value']:.3e} "
            f"atom-steps/s, tuning {out['ensemble']['tuning']}")
    if not a.no_share:
        out["share"] = run_share(a.share_steps, 10)
        log(f"share: {out['share']['ms_per_step']:.4f} ms/step ({out['share']['exchange_ms_per_step']:.3f} ms "
            f"exchange), {out['share']['value']:.3e} atom-steps/s")
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
        log(f"cpu: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
