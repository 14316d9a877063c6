#!/usr/bin/env python3
"""Mechanics-step benchmark on one MI355X: agent-steps/s of the 88.7k-agent 3D workload.

Workload (one "step" = one mechanics timestep of the whole population): 88,700 agents in
a ball of radius 380 um inside a 4000 um box, two cell types 70/30, springs (elastic
0.05, attach 0.01/min, detach 0.001/min, max 12), type B attacking type A (0.005/min,
damage rate 1) and eating dead cells (0.01/min, 1% start dead), random motility 1 um/min,
basement membrane on, dt_mech 0.1 min, FP32 real_t.  Inputs are resident in HBM before
the timed region starts.

Reports one JSON line: whole-step throughput, the HBM roofline of the dominant kernel
(algorithmic bytes per launch / its HIP-event-timed average duration), and the oracle
(CPU restatement of the reference solver, OpenMP) timed on the host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "agent-steps/s at 88.7k-agent 3D; ensemble agent-steps/s + p99 request latency"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD public specification


def kernel_bytes(kernel: str, d: int, r: int, T: int, k: float, s: float) -> float:
    """Algorithmic HBM bytes per agent of one kernel launch: each agent field the kernel
    needs is read once and each output written once; neighbor and spring partner data are
    counted once per agent (the unique data), not once per gather."""
    pos = d * r
    table = {
        # own pos, rel, radius, voxel id; every agent's voxel-sorted record (x,y,z,rel*r) and
        # bucket slot once; voxel start/count; writes: list entries and count
        "NEIGHBORS": pos + 2 * r + 4 + (4 * r + 4) + 8 + 4 * k + 4,
        # forces: pos, vel (rw), radius, R, A, rel, type, affinity row, movable, list;
        # motility: vector (rw), bias direction, bias, speed, persistence, motile flag;
        # membrane strength; springs: count, list, elastic; writes pressure
        "VELOCITY_FUSED": pos + 2 * pos + 4 * r + 4 + T * r + 1 + 4 + 4 * k
        + 2 * pos + pos + 3 * r + 1 + r + 4 + 4 * s + r + r,
        "POSITIONS": 1 + 3 * pos + 3 * pos,
        "BIN_COUNT": pos + 8 + 8,
        "BIN_ORDER": 4 + 8 + 4 + pos + 2 * r + 4 + 4 * r,
        "INTERACT_TENTATIVE": 1 + 4 + 4 * k + 4 + 4 * T * r + r + 4 + 1 + 1 + 4 + 4 + 1,
        "SPRING_COLLECT": 1 + 4 + 4 * k + 4 + r + T * r,
    }
    return float(table.get(kernel, 0.0))


def measured_traffic(tag: str):
    """HBM bytes per launch of kernel `tag` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (copy-calibrated), or None when the summary has no row for it."""
    path = os.path.join(ROOT, "profiles", "mechanics_step_88k_3d_f32_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                cols = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cols) == 6 and cols[1] == tag:
                    return float(cols[5])
    except OSError:
        return None
    return None


def host_cpu_info() -> dict:
    """The host's CPU counts: the CPUs this process may run on, the OS count, and physical
    cores and sockets as lscpu reports them."""
    info = {"affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    try:
        import subprocess

        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        pairs = {tuple(line.split(",")[:2]) for line in out.splitlines() if line and not line.startswith("#")}
        info["physical_cores"] = len(pairs)
        info["sockets"] = len({p[1] for p in pairs})
        model = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        for line in model.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
    except (OSError, ValueError, Exception):  # lscpu missing or unparsable: counts above stand
        pass
    return info


def cpu_step_times(make, threads: int, warmup: int, steps: int, seconds: float):
    """Oracle in OpenMP mode (the reference's mt19937 draws) with `threads` threads: `warmup`
    untimed steps, then up to `steps` steps timed one by one (stopping early after `seconds`).
    Returns (median per-step ms, mean agents over the timed steps, timed steps)."""
    import numpy as np

    from oracle.pyoracle import OracleSim

    mesh, st, params = make()
    ora = OracleSim(mesh, st.real_bytes, st.defs, st.subs, capacity=st.count, seed=2, threads=threads, rng="mt")
    ora.upload(st)
    ora.initialize_mesh()
    ora.time_steps(params, warmup, 0)
    per_step, agents, total = [], [], 0.0
    while len(per_step) < steps and total < seconds * 1000.0:
        n0 = ora.count
        ms = ora.time_steps(params, 0, 1)
        per_step.append(ms)
        agents.append(0.5 * (n0 + ora.count))
        total += ms
    ora.close()
    return float(np.median(per_step)), float(np.mean(agents)), len(per_step)


def cpu_baseline_leg(make, threads: int, steps: int, seconds: float, one_thread_steps: int):
    """BASELINE protocol on the host: 10 warm-up steps, then the median per-step time over
    `steps` steps with `threads` threads, and a 1-thread figure over fewer steps."""
    med, agents, done = cpu_step_times(make, threads, 10, steps, seconds)
    med1, agents1, done1 = cpu_step_times(make, 1, 10, one_thread_steps, seconds)
    return {
        "value": round(agents / (med / 1000.0), 1),
        "unit": "agent-steps/s",
        "cores": threads,
        "median_step_ms": round(med, 3),
        "timed_steps": done,
        "one_thread_value": round(agents1 / (med1 / 1000.0), 1),
        "one_thread_median_step_ms": round(med1, 3),
        "one_thread_timed_steps": done1,
    }


def copy_calibration(mib: int) -> None:
    """A device-to-device copy of ``mib`` MiB, launched 5 times: read mib MiB, write mib MiB."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4}), flush=True)


def pcie_leg(dev, params, steps: int):
    """The drop-in's full host round trip per step (pcm_dropin.hpp upload / sync_to_host):
    count, every field, both link lists and the neighbor lists uploaded from host memory, one
    mechanics step, then every field, both link lists and the neighbor lists downloaded (as
    when host code reads and writes the whole cell_data between steps).  Host buffers are
    ordinary pageable numpy arrays, like the reference's std::vectors."""
    import numpy as np
    from paraphysicell_amd import abi

    host = dev.download()
    bufs = [(abi.FIELD_ID[name], np.ascontiguousarray(host.arrays[name])) for name, _, _ in abi.FIELDS]
    lists = {w: host.links[w] for w in (abi.LINKS_SPRINGS, abi.LINKS_ATTACHED)}
    nbrs = host.neighbors
    field_bytes = sum(b.nbytes for _, b in bufs)

    def i32(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        return a if a.size else np.zeros(1, np.int32)

    n = n0 = dev.count
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dev._call("set_count", dev._h, n)
        for fid, b in bufs:
            dev._call("set_field", dev._h, fid, b.ctypes.data)
        for w, (off, idx) in lists.items():
            o, e = i32(off), i32(idx)  # referenced until the call returns
            dev._call("set_links", dev._h, w, o.ctypes.data, e.ctypes.data)
        o, e = i32(nbrs[0]), i32(nbrs[1])
        dev._call("set_neighbors", dev._h, o.ctypes.data, e.ctypes.data)
        dev.step(params, 1)
        n = dev.count
        for fid, b in bufs:
            dev._call("get_field", dev._h, fid, b.ctypes.data)
        lists = {w: dev.links(w) for w in lists}
        nbrs = dev.neighbors()
    dev.synchronize()
    elapsed = time.perf_counter() - t0
    agent_steps = 0.5 * (n0 + n) * steps
    return {
        "workload": f"the 88.7k-agent step with the drop-in's full host round trip per step ({steps} steps): count, "
                    f"{len(bufs)} fields ({field_bytes / 1e6:.1f} MB), link and neighbor lists up; one step; all down",
        "value": round(agent_steps / elapsed, 1),
        "unit": "agent-steps/s",
        "ms_per_step": round(elapsed * 1000.0 / steps, 3),
        "field_transfer_gbs": round(2 * field_bytes * steps / elapsed / 1e9, 2),
        "note": "PCIe-inclusive rate; not the metric (the timed step keeps the population resident in HBM)",
    }


def ensemble_leg(steps: int):
    """64 replicates (seeds 0-63, 500-20000 agents log-uniform, 2D even / 3D odd seeds, the
    900-agent 2D workload's density and coefficients), `steps` steps each: one batch with
    hipGraph-captured steps against the same simulations launched one at a time."""
    from paraphysicell_amd import synth
    from paraphysicell_amd.serving import run_ensemble

    sizes = synth.config3_sizes()
    members = []
    params = None
    for s_, n in enumerate(sizes):
        mesh, st, params = synth.config3_member(s_, n)
        members.append((mesh, st, s_))
    t_b, as_b, sims = run_ensemble(members, params, steps, batched=True, graphs=True)
    for sim in sims:
        sim.close()
    t_s, as_s, sims = run_ensemble(members, params, steps, batched=False)
    for sim in sims:
        sim.close()
    return {
        "workload": f"{len(sizes)} replicates, 500-20000 agents log-uniform (total {sum(sizes)}), 2D/3D alternating, "
                    f"900-agent-2D density and coefficients, {steps} steps each, FP32",
        "value": round(as_b / t_b, 1),
        "unit": "agent-steps/s",
        "ms_per_step": round(t_b * 1000.0 / steps, 4),
        "sequential_value": round(as_s / t_s, 1),
        "speedup_vs_sequential": round((as_b / t_b) / (as_s / t_s), 2),
        "launch": "one batch, hipGraph-captured step per dims group (grid.y = simulation)",
    }


def serving_leg(count: int, rate: float, steps: int):
    """Poisson arrivals at `rate`/s of `count` requests (1k-200k agents log-uniform, 2D/3D at
    equal odds, the 88.7k-agent workload's physics), `steps` steps each, served by
    continuous batching under a 200 GB HBM budget; the same requests saturating the server
    and run one at a time give the throughput comparison."""
    import numpy as np

    from paraphysicell_amd import synth
    from paraphysicell_amd.serving import Request, Server, poisson_arrivals, run_alone

    arrivals = poisson_arrivals(count, rate, seed=4)
    specs = [synth.serving_request(k, seed=4) for k in range(count)]

    def requests(times):
        return [Request(k, float(times[k]), mesh, st, params, steps, seed=1000 + k)
                for k, (mesh, st, params) in enumerate(specs)]

    # the simulations resident on the device first: batched steps against one at a time,
    # each with the launch shape the tuner picks for it (per group, or per simulation)
    from paraphysicell_amd.serving import run_ensemble

    members = [(mesh, st, 1000 + k) for k, (mesh, st, _) in enumerate(specs)]
    group_shapes, member_shapes = {}, {}
    t_rb, as_rb, sims = run_ensemble(members, specs[0][2], steps, batched=True, shapes=group_shapes)
    for sim in sims:
        sim.close()
    t_rs, as_rs, sims = run_ensemble(members, specs[0][2], steps, batched=False, shapes=member_shapes)
    for sim in sims:
        sim.close()
    # served: the groups' launch shapes tuned above on the same kind of workload
    server = Server(budget_bytes=200e9, chunk=10, graphs=True, tuning=group_shapes)
    done = server.run(requests(arrivals), realtime=True)
    served_rate = server.agent_steps / server.elapsed
    lat = np.array([q.latency for q in done]) * 1000.0
    server.close()
    # one untimed saturated run first: its tear-downs size the pinned read-back pool, as a
    # long-running server's first requests would
    for timed in (False, True):
        sat = Server(budget_bytes=200e9, chunk=10, graphs=True, tuning=group_shapes)
        sat.run(requests(np.zeros(count)), realtime=False)
        sat_rate = sat.agent_steps / sat.elapsed
        sat.close()
    t_alone, as_alone = run_alone(requests(np.zeros(count)), shapes=[member_shapes[k] for k in range(count)])
    alone_rate = as_alone / t_alone
    agents = [st.count for _, st, _ in specs]
    return {
        "workload": f"{count} requests, Poisson {rate}/s, {min(agents)}-{max(agents)} agents (mean "
                    f"{int(np.mean(agents))}), 2D/3D, springs + attack + dead phagocytosis, {steps} steps each, FP32, "
                    "200 GB HBM budget, chunks of 10 steps",
        "throughput": round(served_rate, 1),
        "unit": "agent-steps/s",
        "p50_latency_ms": round(float(np.percentile(lat, 50)), 2),
        "p99_latency_ms": round(float(np.percentile(lat, 99)), 2),
        "saturated_throughput": round(sat_rate, 1),
        "one_at_a_time_throughput": round(alone_rate, 1),
        "speedup_vs_one_at_a_time": round(sat_rate / alone_rate, 2),
        "resident_batched_throughput": round(as_rb / t_rb, 1),
        "resident_one_at_a_time_throughput": round(as_rs / t_rs, 1),
        "resident_speedup": round((as_rb / t_rb) / (as_rs / t_rs), 2),
        "served_launch_shapes": {f"{d}d": list(v) for d, v in sorted(group_shapes.items())},
        "latency": "arrival to result (upload, mesh build, steps, read-back included)",
        "throughput_note": "throughput follows the 20/s arrival rate; saturated_* runs every request at once "
                           "(set-up included; after one untimed saturated run), resident_* steps the same "
                           "simulations already on the device",
    }


ALG_BYTES = {  # algorithmic bytes per agent-step of the whole step (see kernel_bytes for one kernel)
    "disc_900_2d_f32": 120.0,
    "ball_14k_3d_f32": 148.0,
    "ball_14k_3d_f64": 272.0,
    "ball_88k_3d_f32": 252.0,
}


def rocprof_average_us(kernel_prefix: str):
    """Average duration of the kernel whose name starts with `kernel_prefix` in the committed
    rocprofv3 --stats summary of the bench command, or None."""
    path = os.path.join(ROOT, "profiles", "mechanics_step_88k_3d_f32_kernel_stats.md")
    try:
        with open(path) as f:
            for line in f:
                cols = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cols) >= 6 and cols[0].strip("`").startswith(kernel_prefix):
                    return float(cols[2])
    except (OSError, ValueError):
        return None
    return None


def config_leg(name: str, make, steps: int, warmup: int):
    """One workload on its own handle: `warmup` untimed steps, then `steps` timed steps;
    throughput and the whole-step roofline fraction at its algorithmic bytes per agent-step."""
    from paraphysicell_amd.mechanics import DeviceSim

    mesh, st, params = make()
    dev = DeviceSim(mesh, st.real_bytes, st.defs, st.subs, capacity=st.count, seed=2)
    dev.upload(st)
    dev.initialize_mesh()
    dev.step(params, warmup)
    dev.synchronize()
    n0 = dev.count
    t0 = time.perf_counter()
    dev.step(params, steps)
    dev.synchronize()
    elapsed = time.perf_counter() - t0
    n1 = dev.count
    dev.close()
    value = 0.5 * (n0 + n1) * steps / elapsed
    b = ALG_BYTES[name]
    return {
        "agents": st.count,
        "dims": mesh.dims,
        "real_t": "float32" if st.real_bytes == 4 else "float64",
        "steps": steps,
        "value": round(value, 1),
        "unit": "agent-steps/s",
        "ms_per_step": round(elapsed * 1000.0 / steps, 4),
        "alg_bytes_per_agent_step": b,
        "step_roofline_frac": round(value * b / (HBM_PEAK_GBS * 1e9), 5),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600, help="timed mechanics steps (600 = 60 simulated minutes)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--agents", type=int, default=88700)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample length per thread count")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: OMP_NUM_THREADS if set (the host's CPU share for this GPU), else every CPU this "
                         "process may run on")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the 900-agent 2D and 14.3k-agent 3D legs")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-replicate ensemble leg")
    ap.add_argument("--ensemble-steps", type=int, default=600)
    ap.add_argument("--no-serving", action="store_true", help="skip the served-requests leg")
    ap.add_argument("--no-pcie", action="store_true", help="skip the host round-trip leg")
    ap.add_argument("--serving-requests", type=int, default=60)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (5 launches): the reference point that "
                         "scales rocprofv3 FETCH_SIZE/WRITE_SIZE readings to bytes")
    args = ap.parse_args()

    if args.copy_mib:
        copy_calibration(args.copy_mib)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (every simulation here fits one "
              "MI355X; multi-GPU sharding is a design note only).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from paraphysicell_amd import synth
    from paraphysicell_amd.mechanics import DeviceSim, kernel_names

    mesh, st, params = synth.config2(4, n=args.agents)
    dev = DeviceSim(mesh, 4, st.defs, st.subs, capacity=st.count, seed=2)
    dev.upload(st)
    dev.initialize_mesh()

    # warm-up steps: the autotuner picks the launch shape on real steps (1 + 6 * 2 of
    # them) when there are enough of them, then the kernel profile of the last few steps
    # picks the dominant kernel
    autotuned = args.warmup >= 13
    if autotuned:
        dev.autotune(params, 2)
    dev.step(params, max(args.warmup - (13 if autotuned else 0) - 5, 0))
    names = kernel_names()
    dev.profile_begin(names)
    dev.step(params, min(5, args.warmup))
    warm_prof = dev.profile_end()
    dominant = max(warm_prof, key=lambda k: warm_prof[k][0])
    tuned = dev.tuning  # the shape the warm-up steps ran with (tuned, or the library's default for this size)

    n_before = dev.count
    dev.synchronize()
    # the dominant kernel is timed with HIP events on the library's stream over the timed
    # region, every fourth launch: each event pair stalls the stream for a few microseconds,
    # which the step time would carry (about 2 us per step when every second launch is timed)
    dev.profile_every(4)
    dev.profile_begin([dominant])
    t0 = time.perf_counter()
    dev.step(params, args.steps)
    dev.synchronize()
    t1 = time.perf_counter()
    prof = dev.profile_end()
    n_after = dev.count
    elapsed = t1 - t0
    agent_steps = 0.5 * (n_before + n_after) * args.steps
    value = agent_steps / elapsed

    # mean neighbors / springs per agent at the end of the timed region
    off, _ = dev.neighbors()
    k_mean = float(off[-1]) / max(n_after, 1)
    soff, _ = dev.links(0)
    s_mean = float(soff[-1]) / max(n_after, 1)

    ms_total, launches = prof[dominant]
    avg_ms = ms_total / launches
    bytes_per_agent = kernel_bytes(dominant, 3, 4, st.defs, k_mean, s_mean)
    agents_mean = 0.5 * (n_before + n_after)
    bytes_per_launch = bytes_per_agent * agents_mean
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if bytes_per_agent else None
    # the 3D kernel's name carries its wavefronts per voxel, then its candidate tile
    prefix = {"NEIGHBORS": "k_neighbors<float, 3, %d, " % tuned[2]}.get(dominant)
    rp_avg = rocprof_average_us(prefix) if prefix else None
    step_bytes = ALG_BYTES["ball_88k_3d_f32"]

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1000.0 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"3D mechanics step, {args.agents} agents in a 380 um ball, 2 cell types 70/30, springs "
                        "+ attack + dead phagocytosis, BM on, dt_mech 0.1 min",
            "agents": args.agents,
            "agents_at_end": n_after,
            "dims": 3,
            "real_t": "float32",
            "mean_neighbors": round(k_mean, 2),
            "mean_springs": round(s_mean, 2),
            "last_step_counters": dev.stats(),
            "launch_shape": {"velocity_lanes": tuned[0], "event_lanes": tuned[1], "neighbor_waves": tuned[2],
                             "picked_by": "autotuner on warm-up steps" if autotuned
                             else "library default for this population size (fewer than 13 warm-up steps: "
                                  "no tuning)"},
        },
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": measured_traffic(dominant) if args.agents == 88700 else None,
            "traffic_source": "profiles/mechanics_step_88k_3d_f32_hbm_traffic.md (rocprofv3 FETCH_SIZE + "
                              "WRITE_SIZE per launch, copy-calibrated)",
            "bytes_per_agent": round(bytes_per_agent, 1),
            "bytes_per_launch": round(bytes_per_launch, 1),
            "bytes_formula": "kernel_bytes() in bench.py at the end-of-run mean neighbors and springs per agent",
            "avg_launch_us": round(avg_ms * 1000.0, 2),
            "timed_launches": launches,
            "timing": "HIP events on the library stream around every fourth launch of the timed region",
            "rocprof_avg_launch_us": rp_avg,
            "rocprof_source": "profiles/mechanics_step_88k_3d_f32_kernel_stats.md (same command under rocprofv3)",
            "frac_at_rocprof_avg": round(bytes_per_launch / (rp_avg * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
            if rp_avg else None,
            "step_alg_bytes_per_agent": step_bytes,
            "step_frac": round(value * step_bytes / (HBM_PEAK_GBS * 1e9), 5),
        },
        "kernel_share_warmup": {k: round(v[0] / max(sum(x[0] for x in warm_prof.values()), 1e-9), 3)
                                for k, v in sorted(warm_prof.items(), key=lambda kv: -kv[1][0])[:8]},
    }

    if not args.no_pcie:
        result["pcie_inclusive"] = pcie_leg(dev, params, 20)
    dev.close()
    if not args.no_configs:
        result["configs"] = {
            "disc_900_2d_f32": config_leg("disc_900_2d_f32", lambda: synth.config0(4), 600, 10),
            "ball_14k_3d_f32": config_leg("ball_14k_3d_f32", lambda: synth.config1(4), 600, 10),
            "ball_14k_3d_f64": config_leg("ball_14k_3d_f64", lambda: synth.config1(8), 600, 10),
        }
    if not args.no_ensemble:
        result["ensemble"] = ensemble_leg(args.ensemble_steps)
    if not args.no_serving:
        result["serving"] = serving_leg(args.serving_requests, 20.0, 100)
    if not args.no_cpu:
        # the host's CPU share for this GPU (OMP_NUM_THREADS, set by the GPU host's
        # environment), else every CPU this process may run on; OpenMP places one thread
        # per core, as BASELINE's protocol asks, before the oracle's runtime starts
        env_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        info = host_cpu_info()
        threads = args.cpu_threads or env_threads or info["affinity_cpus"]
        os.environ.setdefault("OMP_PROC_BIND", "close")
        os.environ.setdefault("OMP_PLACES", "cores")
        cpu = cpu_baseline_leg(lambda: synth.config2(4, n=args.agents), threads, 100, args.cpu_seconds, 20)
        cpu.update({
            "kind": "port",
            "sample": f"the same workload from its initial state: 10 warm-up steps, then the median of "
                      f"{cpu['timed_steps']} single-step timings with {threads} threads and of "
                      f"{cpu['one_thread_timed_steps']} with 1 thread; the OpenMP oracle with the reference's "
                      "mt19937 draws, built -O3 -fopenmp",
            "host": info,
            "threads_source": "--cpu-threads" if args.cpu_threads else (
                "OMP_NUM_THREADS (the host's CPU share for one GPU)" if env_threads else "sched_getaffinity"),
            "gpu_over_cpu": round(value / cpu["value"], 1),
            "gpu_over_one_thread": round(value / cpu["one_thread_value"], 1),
            "target_30x_met": value / cpu["value"] >= 30.0,
        })
        result["cpu_baseline"] = cpu
    print(json.dumps(result), flush=True)
    del np


if __name__ == "__main__":
    main()
