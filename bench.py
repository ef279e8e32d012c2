"""Benchmark: stochastic LLG-Heun spin-steps/s at 4M atoms (FCC 100^3, 12 neighbours, 600 K),
FP32 fused kernels (the headline value) beside FP64 fused and the plain FP64 kernels, with the
HBM roofline of the Heun step and the reference's own CPU Heun step timed on the host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--L 100] [--cpu-seconds 10] [--no-cpu]
                    [--batching-steps 200] [--legs fp32_fused,fp64_fused,fp64_plain,fp32_midpoint]
                    [--mc-sweeps 200] [--configs c0,c3,c4]

One "step" is one Heun step (predictor + corrector) over all 4,000,000 spins, inputs resident
in HBM.  Also reported: the FP32-vs-FP64 drift of the magnetisation, checkerboard and
constrained Metropolis on the 256-replica batch, the replica-batching leg (256 SC replicas in one
context against one context per replica), the other workloads (C0, the C3 adjoint, the C4 film),
and on the host the
reference's sim::LLG_Heun (the drop-in executable with gpu:acceleration=off, one thread) with the
oracle beside it.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "spin-steps/sec LLG-Heun at 4M atoms, FP32 vs FP64, % of HBM roofline"
HBM_PEAK_GBS = 8000.0


def b_alg(p: int, z: int, c_j: int = 0) -> int:
    """Algorithmic bytes per spin per Heun step: read S, write S', read S', write S (3p each),
    z int32 neighbour indices, c_J exchange reals per edge, one int32 material id."""
    return 12 * p + z * (4 + c_j * p) + 4


def run_gpu(system, precision, kernel, steps, warmup, seed=7, T=600.0, use_graph=False, block=0, integrator="heun"):
    from vampire_amd.context import Context
    spins = np.zeros((3, system.n))
    spins[2] = 1.0
    c = Context(system, precision=precision, kernel=kernel, seed=seed, block_size=max(block, 0), device=0)
    c.set_conditions(T)
    if block < 0:                 # autotune block size and XCD mapping (outside the timed region)
        c.set_spins(spins)
        c.autotune(20)
    c.set_spins(spins)
    step = c.llg_midpoint if integrator == "midpoint" else c.llg_heun
    if warmup:
        step(warmup)
    c.synchronize()
    t0 = time.perf_counter()
    if integrator == "midpoint":
        step(steps)
    elif use_graph:
        chunk = 50
        c.llg_heun_graph(chunk, steps // chunk)
        rest = steps - chunk * (steps // chunk)
        if rest:
            c.llg_heun(rest)
    else:
        c.llg_heun(steps)
    c.synchronize()
    wall = time.perf_counter() - t0
    # HIP events recorded on the kernels' stream before and after the timed steps
    c.event_ms_per_step = c.last_step_time_ms()
    # per-kernel HIP-event timings of the Heun step (the midpoint leg reports its wall time only)
    kernel_ms = c.time_kernels(20) if integrator == "heun" else np.zeros(0)
    c.bench_launch = c.launch_config
    return c, wall, kernel_ms


def batching_leg(steps: int, graph_chunk: int = 50):
    """Config C2: 256 SC replicas (8-48 cells per edge, 8.67M atoms, each at its own temperature in
    [0, 1.5 Tc], alpha = 1, FP32).  One context integrating every replica in the same launches
    (eager and HIP-graph replay) against one context per replica run one after the other, and the
    checkerboard Metropolis sweep the same two ways.  Spin-steps/s over all replicas."""
    from vampire_amd import system as S
    from vampire_amd.context import FP32, FUSED, Context
    reps, temps = S.config_c2()
    total = sum(r.n for r in reps)
    batched = S.concat_systems(reps)
    out = {"workload": f"{len(reps)} SC replicas, L in [8,48], {total} atoms, T uniform in [0, 1.5 Tc], "
                       "alpha=1, FP32 fused LLG-Heun; checkerboard Metropolis",
           "replicas": len(reps), "atoms": total, "steps": steps}

    def up(n):
        sp = np.zeros((3, n))
        sp[2] = 1.0
        return sp

    c = Context(batched, precision=FP32, kernel=FUSED, seed=7, device=0)
    for r, T in enumerate(temps):
        c.set_conditions(T, replica=r)
    c.set_spins(up(total))
    c.llg_heun(20)
    c.synchronize()
    t0 = time.perf_counter()
    c.llg_heun(steps)
    c.synchronize()
    out["batched_sps"] = total * steps / (time.perf_counter() - t0)
    c.llg_heun_graph(graph_chunk, 1)
    c.synchronize()
    t0 = time.perf_counter()
    c.llg_heun_graph(graph_chunk, steps // graph_chunk)
    c.synchronize()
    out["batched_graph_sps"] = total * graph_chunk * (steps // graph_chunk) / (time.perf_counter() - t0)
    c.mc_sweep(2)
    c.synchronize()
    t0 = time.perf_counter()
    c.mc_sweep(steps)
    c.synchronize()
    out["mc_batched_sps"] = total * steps / (time.perf_counter() - t0)
    c.close()

    ctxs = []
    for r, (s_r, T) in enumerate(zip(reps, temps)):
        cr = Context(s_r, precision=FP32, kernel=FUSED, seed=7 + r, device=0)
        cr.set_conditions(T)
        cr.set_spins(up(s_r.n))
        cr.llg_heun(2)
        cr.mc_sweep(1)
        cr.synchronize()
        ctxs.append(cr)
    t0 = time.perf_counter()
    for cr in ctxs:                 # one replica at a time, its launches back to back
        cr.llg_heun(steps)
        cr.synchronize()
    out["sequential_sps"] = total * steps / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for cr in ctxs:
        cr.mc_sweep(steps)
        cr.synchronize()
    out["mc_sequential_sps"] = total * steps / (time.perf_counter() - t0)
    for cr in ctxs:
        cr.close()
    out["speedup"] = round(max(out["batched_sps"], out["batched_graph_sps"]) / out["sequential_sps"], 2)
    out["mc_speedup"] = round(out["mc_batched_sps"] / out["mc_sequential_sps"], 2)
    for k in ("batched_sps", "batched_graph_sps", "sequential_sps", "mc_batched_sps", "mc_sequential_sps"):
        out[k] = round(out[k], 1)
    return out


MC_BYTES_PER_MOVE = 64     # SC checkerboard Metropolis, FP32: 9p + 4z + 4 with z = 6


def metropolis_leg(sweeps: int, warmup: int = 5):
    """Config C2 under checkerboard Metropolis: the 256-replica batch (8.67M atoms, each replica at
    its own temperature) swept as one context, FP32 and FP64, spins from +z; and the constrained
    Monte Carlo sweep of the same batch (FP32, every replica constrained at phi = 40, theta = 30
    degrees).  Spin-moves (pair trials for constrained MC: N per sweep) per second over all
    replicas over the wall-clock window of one call of `sweeps` sweeps (the call moves the spins to
    colour order and back once), with the FP32 sweep against 64 B per spin-move at 8 TB/s."""
    from vampire_amd import system as S
    from vampire_amd.context import FP32, FP64, Context
    reps, temps = S.config_c2()
    batched = S.concat_systems(reps)
    total = batched.n
    out = {"workload": f"{len(reps)} SC replicas, L in [8,48], {total} atoms, T uniform in [0, 1.5 Tc], adaptive-cone "
                       "checkerboard Metropolis from +z",
           "atoms": total, "sweeps": sweeps, "bytes_per_move": MC_BYTES_PER_MOVE}
    for prec in (FP32, FP64):
        c = Context(batched, precision=prec, seed=7, device=0)
        for r, T in enumerate(temps):
            c.set_conditions(T, replica=r)
        sp = np.zeros((3, total))
        sp[2] = 1.0
        c.set_spins(sp)
        c.mc_sweep(warmup)
        c.synchronize()
        t0 = time.perf_counter()
        c.mc_sweep(sweeps)
        c.synchronize()
        wall = time.perf_counter() - t0
        key = "fp32" if prec == FP32 else "fp64"
        out[key] = {"value": round(total * sweeps / wall, 1), "ms_per_sweep": round(1e3 * wall / sweeps, 5),
                    "colours": c.mc_num_colours()}
        if prec == FP32:
            achieved = MC_BYTES_PER_MOVE * total * sweeps / wall / 1e9
            out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(achieved / HBM_PEAK_GBS, 4)}
            c.cmc_set_constraint(40.0, 30.0)
            c.set_spins(sp)
            c.cmc_sweep(warmup)
            c.synchronize()
            t0 = time.perf_counter()
            c.cmc_sweep(sweeps)
            c.synchronize()
            wall = time.perf_counter() - t0
            out["cmc_fp32"] = {"value": round(total * sweeps / wall, 1), "ms_per_sweep": round(1e3 * wall / sweeps, 5),
                               "unit": "pair trials/s"}
        c.close()
    return out


def configs_leg(c0_steps: int = 10000, c3_steps: int = 2000, c4_steps: int = 1000, only=("c0", "c3", "c4")):
    """The other workloads of the path, timed on the device:
    C0  SC 20^3 (8000 atoms), 300 K, FP64 fused, HIP-graph replay at the autotuned launch shape
        (8000 atoms fill few CUs, so the shape matters); the oracle on one host thread for 1000 steps
        beside it (the reference's own CPU-runnable case);
    C3  BCC 32^3 (65,536 atoms), deterministic Heun under Hz = 1 T from a 30-degree tilt, 2000 steps,
        checkpoint every 100: trajectory forward and hand-written adjoint backward, FP32 and FP64;
    C4  ferrimagnetic SC film 256x256x8 with interfacial DMI (3x3 tensor exchange), 10% dilution
        (472,014 atoms), FP32 fused, temperature set every step along a Gaussian 100 fs pulse
        300 -> 1200 K (each step: vmx_set_conditions + one Heun step, as gpu::llg_heun does)."""
    from vampire_amd import system as S
    from vampire_amd.context import FP32, FP64, FUSED, Context
    from vampire_amd.fit import Trajectory
    out = {}
    if "c0" in only:
        out["c0"] = _c0_leg(S, FP64, FUSED, Context, c0_steps)
    if "c3" in only:
        out["c3"] = _c3_leg(S, FP32, FP64, Trajectory, c3_steps)
    if "c4" in only:
        out["c4"] = _c4_leg(S, FP32, FUSED, Context, c4_steps)
    return out


def _c0_leg(S, FP64, FUSED, Context, c0_steps):
    s0 = S.config_c0()
    sp0 = S.random_spins(s0.n, 1)
    c = Context(s0, precision=FP64, kernel=FUSED, seed=1, device=0)
    c.set_conditions(300.0)
    c.set_spins(sp0)
    c.autotune(20)                # launch shape for this system, outside the timed region
    c.set_spins(sp0)
    c.llg_heun_graph(100, 1)
    c.synchronize()
    t0 = time.perf_counter()
    c.llg_heun_graph(100, c0_steps // 100)
    c.synchronize()
    dt = time.perf_counter() - t0
    launch = c.launch_config
    c.close()
    from oracle import oracle as O
    o = O.Oracle(s0)
    o.set_conditions(300.0)
    t1 = time.perf_counter()
    o.llg_heun(sp0, 1000, seed=1, threads=1)
    cpu = s0.n * 1000 / (time.perf_counter() - t1)
    return {"workload": "SC 20^3, 8000 atoms, 300 K, FP64 fused, graph replay", "steps": c0_steps,
            "seconds": round(dt, 4), "spin_steps_per_s": round(s0.n * c0_steps / dt, 1),
            "launch": {"block": launch[0], "xcd_remap": launch[1]},
            "cpu_1thread_spin_steps_per_s": round(cpu, 1)}


def _c3_leg(S, FP32, FP64, Trajectory, c3_steps):
    s3 = S.config_c3()
    th = np.deg2rad(30.0)
    sp3 = np.zeros((3, s3.n))
    sp3[0], sp3[2] = np.sin(th), np.cos(th)
    rng = np.random.default_rng(3)
    h = 0.01 * rng.standard_normal((3, s3.n))
    c3 = {"workload": "BCC 32^3, 65536 atoms, T=0, Hz=1 T, 2000 steps, checkpoint 100, adjoint in J, Ku, alpha, h",
          "steps": c3_steps}
    for prec in (FP32, FP64):
        t = Trajectory(s3, sp3, nsteps=c3_steps, record_every=10, checkpoint_every=100, precision=prec,
                       H=(0.0, 0.0, 1.0))
        mz = t.forward(S.J_CO, S.KU_CO, 0.1, h)
        t.backward(np.ones_like(mz))
        ta = time.perf_counter()
        mz = t.forward(S.J_CO * 1.1, S.KU_CO * 0.9, 0.12, h)
        tb = time.perf_counter()
        t.backward(np.linspace(-1.0, 1.0, mz.size))
        tc = time.perf_counter()
        t.close()
        c3[prec] = {"forward_s": round(tb - ta, 4), "backward_s": round(tc - tb, 4),
                    "forward_spin_steps_per_s": round(s3.n * c3_steps / (tb - ta), 1)}
    return c3


def _c4_leg(S, FP32, FUSED, Context, c4_steps):
    s4 = S.config_c4()
    c = Context(s4, precision=FP32, kernel=FUSED, seed=3, device=0)
    c.set_spins(S.random_spins(s4.n, 3))
    name, npat = c.topology
    tau, t_peak = 100e-15, 0.5 * c4_steps * 1e-16
    temps = [300.0 + 900.0 * np.exp(-((k * 1e-16 - t_peak) / tau) ** 2) for k in range(c4_steps)]
    c.set_conditions(300.0)
    c.llg_heun(20)
    c.synchronize()
    t0 = time.perf_counter()
    for T in temps:
        c.set_conditions(T)
        c.llg_heun(1)
    c.synchronize()
    dt = time.perf_counter() - t0
    k4 = c.time_kernels(20)
    launch = c.launch_config
    c.close()
    return {"workload": "SC film 256x256x8, DMI tensor exchange, 10% dilution, 472014 atoms, FP32 fused, "
                        "per-step Gaussian T pulse 300->1200 K", "atoms": s4.n, "steps": c4_steps,
            "topology": name, "stencils": npat, "seconds": round(dt, 4),
            "spin_steps_per_s": round(s4.n * c4_steps / dt, 1), "kernel_ms": k4.tolist(),
            "launch": {"block": launch[0], "xcd_remap": launch[1]}}


def drift_leg(system, steps: int, c0_steps: int = 1000):
    """FP32 against FP64 after `steps` Heun steps of the benchmark system from the same state (600 K,
    spins +z, seed 7): with the same normals (the FP32 context draws the FP64 generator's normals,
    fp64_normals) the difference is the FP32 arithmetic alone; with its own FP32 normals it also
    holds the generators' difference.  max_abs_dm is the largest per-spin-component difference,
    dM_over_Ms = |<m>_FP32 - <m>_FP64|.  The trajectories are chaotic, so per-spin differences
    grow from any rounding: c0_floor measures, on the SC 20^3 workload at 300 K over c0_steps, the
    same-normals FP32 difference beside what FP64 arithmetic reaches when its spins are rounded to
    float after every step, and what one 2^-24 perturbation of the FP64 start grows into."""
    from vampire_amd import system as S
    from vampire_amd.context import FP32, FP64, FUSED, Context

    def run(sys_, prec, f64n, T, start, k, seed, store32=False):
        c = Context(sys_, precision=prec, kernel=FUSED, seed=seed, device=0, fp64_normals=f64n)
        c.set_conditions(T)
        c.set_spins(start)
        if store32:
            for _ in range(k):
                c.llg_heun(1)
                c.set_spins(c.get_spins().astype(np.float32).astype(np.float64))
        else:
            c.llg_heun(k)
        out = c.get_spins()
        c.close()
        return out

    def cmp(a, b):
        return dict(max_abs_dm=float(np.abs(a - b).max()),
                    dM_over_Ms=float(np.linalg.norm(a.mean(axis=1) - b.mean(axis=1))))

    sp = np.zeros((3, system.n))
    sp[2] = 1.0
    r64 = run(system, FP64, False, 600.0, sp, steps, 7)
    same = cmp(run(system, FP32, True, 600.0, sp, steps, 7), r64)
    own = cmp(run(system, FP32, False, 600.0, sp, steps, 7), r64)
    s0 = S.config_c0()
    sp0 = S.random_spins(s0.n, 3)
    rng = np.random.default_rng(1)
    pert = sp0 * (1.0 + rng.choice([-1.0, 1.0], sp0.shape) * 2.0 ** -24)
    pert /= np.linalg.norm(pert, axis=0)
    c64 = run(s0, FP64, False, 300.0, sp0, c0_steps, 9)
    floor = dict(workload=f"SC 20^3, 8000 atoms, 300 K, {c0_steps} steps",
                 fp32_same_normals=cmp(run(s0, FP32, True, 300.0, sp0, c0_steps, 9), c64)["max_abs_dm"],
                 fp64_stored_as_float=cmp(run(s0, FP64, False, 300.0, sp0.astype(np.float32).astype(np.float64),
                                              c0_steps, 9, store32=True), c64)["max_abs_dm"],
                 fp64_one_perturbation=cmp(run(s0, FP64, False, 300.0, pert, c0_steps, 9), c64)["max_abs_dm"])
    return dict(steps=steps, same_normals=same, own_normals=own, target_max_abs_dm=1e-5, c0_floor=floor,
                max_abs_dm=same["max_abs_dm"], dM_over_Ms=same["dM_over_Ms"])


def host_cpu():
    """(model name, threads the host gives this process) for the CPU baseline: the cores in the
    process's affinity mask, capped by OMP_NUM_THREADS where the environment sets the host share."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        cores = min(cores, int(env))
    return model, cores


def cpu_sample(system, threads, seconds):
    """Oracle Heun steps on the host until the sample lasts `seconds`: (spin-steps/s, steps)."""
    one = cpu_baseline(system, 1, threads)
    steps = max(1, int(round(seconds * one / system.n)))
    return cpu_baseline(system, steps, threads), steps


REF_EXE = os.path.join(ROOT, "vampire_amd", "shim", "_build", "vampire-vmx")
# The CPU-baseline workload as a VAMPIRE input deck: FCC 100^3 cells (4M atoms), periodic,
# J = 6.064e-21 J, Ku = 4.644e-24 J along z, mu_s = 1.72 mu_B, alpha = 0.05, spins +z, 600 K,
# dt = 1e-16 s, sim:program=benchmark (src/program/bmark.cpp: the steps, then one statistics
# update and one data line).
REF_INPUT = """create:crystal-structure=fcc
create:periodic-boundaries-x
create:periodic-boundaries-y
create:periodic-boundaries-z
dimensions:unit-cell-size = 3.54 !A
dimensions:system-size-x = {size} !A
dimensions:system-size-y = {size} !A
dimensions:system-size-z = {size} !A
material:file=Co.mat
sim:temperature=600
sim:time-steps-increment={steps}
sim:total-time-steps={steps}
sim:time-step=1.0E-16
sim:program=benchmark
sim:integrator=llg-heun
gpu:acceleration=off
output:time-steps
"""
REF_MAT = """material:num-materials=1
material[1]:material-name=Co
material[1]:damping-constant=0.05
material[1]:exchange-matrix[1]=6.064e-21
material[1]:atomic-spin-moment=1.72 !muB
material[1]:second-order-uniaxial-anisotropy-constant=4.644e-24
material[1]:initial-spin-direction=0,0,1
"""


def reference_seconds(cells, steps):
    """Wall time of one run of the reference's own CPU integrator on FCC cells^3: the drop-in
    executable (VAMPIRE built from its sources, its src/gpu dispatchers replaced) with
    gpu:acceleration=off runs sim::LLG_Heun on one thread, as vampire-serial does.  Set-up
    included."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "input"), "w") as f:
            f.write(REF_INPUT.format(size=f"{3.54 * cells - 0.01:.2f}", steps=steps))
        with open(os.path.join(d, "Co.mat"), "w") as f:
            f.write(REF_MAT)
        t0 = time.perf_counter()
        subprocess.run([REF_EXE], cwd=d, capture_output=True, timeout=900, check=True)
        dt = time.perf_counter() - t0
        rows = [ln for ln in open(os.path.join(d, "output")).read().splitlines() if ln.strip() and ln[0] != "#"]
        assert rows and int(float(rows[-1].split()[0])) == steps, "reference run ended early"
        return dt


def reference_rate(n_atoms, cells, steps):
    """Spin-steps/s of the reference's Heun step: the difference of a 2-step and a
    (2 + steps)-step run, which cancels the set-up."""
    t_short = reference_seconds(cells, 2)
    t_long = reference_seconds(cells, 2 + steps)
    return n_atoms * steps / max(t_long - t_short, 1e-9)


def cpu_baseline(system, nsteps, threads):
    from oracle import oracle as O
    o = O.Oracle(system)
    o.set_conditions(600.0)
    spins = np.zeros((3, system.n))
    spins[2] = 1.0
    o.llg_heun(spins, 1, seed=7, threads=threads)       # first touch
    t0 = time.perf_counter()
    o.llg_heun(spins, nsteps, seed=7, threads=threads)
    return system.n * nsteps / (time.perf_counter() - t0)


def copy_calibration(nbytes: int = 256 << 20, reps: int = 3) -> None:
    """Device-to-device copies of a known size on the HIP runtime libvmx uses: under
    rocprofv3 --pmc their FETCH_SIZE / WRITE_SIZE give the counters' scale in bytes."""
    import ctypes as C
    from vampire_amd import _lib
    _lib.load()
    hip = C.CDLL("libamdhip64.so.7")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    for _ in range(reps):
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"# copy calibration: {reps} x {nbytes >> 20} MiB device-to-device", file=sys.stderr, flush=True)


def measured_traffic(n: int):
    """HBM bytes per Heun step from the committed rocprofv3 --pmc summary of this workload
    (FETCH_SIZE + WRITE_SIZE of both kernels, scaled by the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "fcc100_fp32_fused_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("atoms") != n:
        return None
    return d.get("traffic_bytes_per_step")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--L", type=int, default=100, help="FCC unit cells per edge (100 -> 4M atoms)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU work per thread count: steps are added until the sample lasts this long")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay the fused step as a HIP graph")
    ap.add_argument("--drift-steps", type=int, default=1000)
    ap.add_argument("--block", type=int, default=-1, help="threads per workgroup; 0: library default, -1: autotune")
    ap.add_argument("--batching-steps", type=int, default=200, help="0 skips the replica-batching leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the C0 / C3 / C4 leg")
    ap.add_argument("--configs", default="c0,c3,c4", help="comma list of the workloads the configs leg runs")
    ap.add_argument("--legs", default="fp32_fused,fp64_fused,fp64_plain,fp32_midpoint",
                    help="comma list of fp32_fused, fp64_fused, fp64_plain, fp32_midpoint, fp64_midpoint "
                         "(profiling runs use one)")
    ap.add_argument("--mc-sweeps", type=int, default=200, help="0 skips the Metropolis leg (C2 batch)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with three 256 MiB device-to-device copies (counter scale for traffic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this build runs on one GPU "
              "(replicas and spatial z-slabs with a halo exchange would run one process per GPU)", file=sys.stderr)
        sys.exit(2)

    from vampire_amd import system as S
    from vampire_amd.context import FP32, FP64, FUSED, PLAIN

    t_build = time.perf_counter()
    system = S.config_c1(args.L)
    t_build = time.perf_counter() - t_build
    n = system.n
    z = int(system.nbr_end[0] - system.nbr_start[0] + 1)
    print(f"# system: FCC {args.L}^3, {n} atoms, z={z}, built in {t_build:.1f} s", file=sys.stderr, flush=True)

    legs = [x for x in args.legs.split(",") if x]
    table = {"fp32_fused": (FP32, FUSED, "heun"), "fp64_fused": (FP64, FUSED, "heun"),
             "fp64_plain": (FP64, PLAIN, "heun"), "fp32_midpoint": (FP32, FUSED, "midpoint"),
             "fp64_midpoint": (FP64, FUSED, "midpoint")}
    res = {}
    for name in legs:
        prec, kern, integ = table[name]
        c, wall, kms = run_gpu(system, prec, kern, args.steps, args.warmup, use_graph=args.graph and kern == FUSED,
                               block=args.block, integrator=integ)
        res[name] = dict(wall=wall, kernel_ms=kms.tolist(), sps=n * args.steps / wall,
                         ms_per_step=1e3 * wall / args.steps, event_ms_per_step=c.event_ms_per_step,
                         launch={"block": c.bench_launch[0], "xcd_remap": c.bench_launch[1]})
        c.close()
        print(f"# {name}: {res[name]['ms_per_step']:.4f} ms/step, {res[name]['sps']:.3e} spin-steps/s, "
              f"kernels {['%.4f' % v for v in kms]} ms, launch {res[name]['launch']}", file=sys.stderr, flush=True)
    if args.copy_calibration:
        copy_calibration()

    drift = drift_leg(system, args.drift_steps) if args.drift_steps else {}
    if drift:
        print(f"# drift: {drift}", file=sys.stderr, flush=True)

    head = res.get("fp32_fused")
    roofline = None
    if head is not None:
        p32 = b_alg(4, z)
        # the two kernels of a Heun step as one unit: their duration per step from the stream
        # events around the timed region (falls back to the per-kernel event timings)
        step_ms = head["event_ms_per_step"] if head["event_ms_per_step"] > 0 else float(sum(head["kernel_ms"]))
        achieved = p32 * n / (step_ms * 1e-3) / 1e9
        roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(n),
                        kernel="fused predictor + corrector (one Heun step)",
                        bytes_per_spin_step=p32, step_ms=round(step_ms, 5), kernel_ms=head["kernel_ms"],
                        launch=head["launch"])
    metropolis = None
    if args.mc_sweeps:
        metropolis = metropolis_leg(args.mc_sweeps)
        print(f"# metropolis: {metropolis}", file=sys.stderr, flush=True)

    batching = None
    if args.batching_steps:
        batching = batching_leg(args.batching_steps)
        print(f"# batching: {batching}", file=sys.stderr, flush=True)

    configs = None
    if not args.no_configs:
        configs = configs_leg(only=[x for x in args.configs.split(",") if x])
        print(f"# configs: {configs}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        model, nthreads = host_cpu()
        (v1, k1), (vN, kN) = (cpu_sample(system, 1, args.cpu_seconds),
                              cpu_sample(system, nthreads, args.cpu_seconds))
        port = dict(value=round(vN, 1), cores=nthreads, steps=kN, value_1thread=round(v1, 1), steps_1thread=k1,
                    cpu_model=model, host_cpus=os.cpu_count(),
                    what="FP64 oracle (the reference's loops restated in C), same workload, OpenMP over every "
                         "core the host gives this process")
        print(f"# cpu port: 1 thread {v1:.3e}, {nthreads} threads {vN:.3e} spin-steps/s", file=sys.stderr, flush=True)
        vr, why = None, "drop-in executable not built (make -C vampire_amd/shim)"
        ref_steps = max(4, int(round(args.cpu_seconds * 2.5)))
        if os.path.exists(REF_EXE):
            try:
                vr = reference_rate(system.n, args.L, ref_steps)
            except Exception as e:       # the port stays the baseline; the reason goes in the line
                vr, why = None, f"reference run failed: {type(e).__name__}: {str(e)[:200]}"
        if vr is not None:
            cpu = dict(value=round(vr, 1), unit="spin-steps/s", cores=1, kind="reference",
                       sample=f"{ref_steps} Heun steps of the same FCC {args.L}^3 system ({system.n} atoms, 600 K, spins +z) by the "
                              f"reference's own sim::LLG_Heun: VAMPIRE built from its sources with gpu:acceleration=off, "
                              f"one thread as vampire-serial; the rate is the difference of a 2-step and a "
                              f"{2 + ref_steps}-step run, which cancels the set-up",
                       cpu_model=model, port=port, port_1thread_over_reference=round(v1 / vr, 3))
            print(f"# cpu reference: {vr:.3e} spin-steps/s (1 thread)", file=sys.stderr, flush=True)
        else:
            cpu = dict(value=port["value"], unit="spin-steps/s", cores=nthreads, kind="port", cpu_model=model,
                       sample=f"{kN} Heun steps of the same 4M-atom FCC system (FP64 oracle, OpenMP over {nthreads} "
                              f"threads, >= {args.cpu_seconds:.0f} s); 1 thread: {k1} steps, {v1:.4g} spin-steps/s",
                       value_1thread=port["value_1thread"], reference_skipped=why)

    def leg(name, p):
        if name not in res:
            return None
        r = res[name]
        return {"value": round(r["sps"], 1), "ms_per_step": round(r["ms_per_step"], 5), "kernel_ms": r["kernel_ms"],
                "launch": r["launch"],
                "hbm_frac_b_alg": round(b_alg(p, z) * n / (sum(r["kernel_ms"]) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    out = {
        "metric": METRIC,
        "value": None if head is None else round(head["sps"], 1),
        "unit": "spin-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None if head is None else round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"FCC {args.L}^3 ferromagnet, {n} atoms, 12 NN scalar J, Ku along z, alpha=0.05, "
                               "T=600 K, spins +z, Philox seed 7, stochastic LLG-Heun dt=1e-16 s",
                   "atoms": n, "neighbours": z, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp64_fused": leg("fp64_fused", 8),
        "fp64_plain": leg("fp64_plain", 8),
        "fp32_over_fp64_plain": (round(head["sps"] / res["fp64_plain"]["sps"], 3)
                                 if head is not None and "fp64_plain" in res else None),
        "fp32_midpoint": ({"value": round(res["fp32_midpoint"]["sps"], 1),
                           "ms_per_step": round(res["fp32_midpoint"]["ms_per_step"], 5),
                           "integrator": "LLG midpoint (sim::LLG_Midpoint), same workload"}
                          if "fp32_midpoint" in res else None),
        "fp32_vs_fp64_drift": drift,
        "metropolis": metropolis,
        "batching": batching,
        "configs": configs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
