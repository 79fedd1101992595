#!/usr/bin/env python3
"""Benchmark: syn events/s and x-realtime of the 1M-neuron / 100M-synapse Izhikevich + pair-STDP
network on one MI355X, with the dominant kernel's roofline and a CPU baseline; the other
workloads (10k LIF, 200k HH plain vs surrogate, the 256-network population, 10^8 LIF) as
sections of the same JSON line.

A bench step is one SimulationEngine::simulate_ms(100) call (engine.rs:792-797): 100 simulation
steps of 1 ms of the whole network (neuron update, propagation with delays, STDP), so the default
100 steps are the workload's 10 s of simulated time.  The network's activity oscillates
(population bursts about every 0.2 s), and a 100 ms step keeps any timed window several cycles
long.  Before any timing the network runs a declared burn-in (`config.burnin_ms`, default 1 s
simulated) out of its quiet initial state, then the --warmup steps.  `value` = synaptic events
(sum of out-degrees of spiking rows) per wall second over exactly the --steps timed steps, network
and state resident in HBM.  The same steps are then replayed from a device snapshot as direct
launches bracketed by hipEvents, so the per-kernel times and the roofline describe the timed
window itself.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
BF16_PEAK = 2.5e15     # MI355X dense bf16 MFMA, FLOP/s
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "traffic_izh_stdp_1M_100M.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def neuron_bytes(cfg, n, spikes, ring=True):
    """Algorithmic bytes of one update launch, a lower bound: per-neuron state read and the
    writes that always happen, ring slots read, plus per-spike state (jitter c/d, count, last
    step, pre trace record, spike list).  Writes the kernel skips when the value's bits do not
    change (a ring slot that received nothing, LIF currents that stay zero, a V at rest) are not
    counted.  ring=False: no spike in the last D - 1 steps, so no slot can hold input and the
    update does not read the ring."""
    from neuralsim_amd import abi
    if cfg.model == abi.MODEL_IZHIKEVICH:
        per_n, per_s, ring_b = 8 + 8 + 8, 8 + 8 + 4 + 16 + 4, 8  # v, u read+write; ring read
    elif cfg.model == abi.MODEL_LIF:
        per_n, per_s, ring_b = 16 + 16 + 4, 8 + 4 + 4, 16  # v, 2 currents, counter, 2 slots; v
    else:
        per_n, per_s, ring_b = 4 * 8 * 2 + 16, 8 + 4 + 4, 16  # 8 state words r+w; 2 slots
    return (per_n - (0 if ring else ring_b)) * n + per_s * spikes


def synapse_bytes(events, spikes, ltp_slots, ltd_slots):
    """Algorithmic bytes of the synapse pass: propagation 4 B col + 4 B weight + 1 B delay +
    8 B int64 accumulate per event and 16 B of row pointers per spiking row; potentiation 16 B
    post record + 4 B weight write per E->E out-slot of a spiking row; depression 4 B slot + 4 B
    source + 8 B pre record + 8 B weight read+write per walked E->E in-slot."""
    prop = 17.0 * events + 16.0 * spikes
    return prop, prop + 20.0 * ltp_slots + 24.0 * ltd_slots


def regime_tag(args):
    return {"workload": "izhikevich_pair_stdp_1M_100M", "neurons": args.neurons,
            "burnin_ms": args.burnin_ms, "step_ms": args.step_ms, "warmup": args.warmup,
            "steps": args.steps, "delivery": args.delivery}


def set_headline_delivery(ns, eng, delivery):
    """Delivery of the headline network: binned (the default, measured faster on this network)
    or ring atomics; "auto" leaves it to the autotuner, run by the caller."""
    if delivery == "binned":
        eng.set_delivery(ns.abi.DELIVERY_BINNED, 0, 0)
    elif delivery == "atomic":
        eng.set_delivery(ns.abi.DELIVERY_ATOMIC)


def measured_traffic(kernel, tag):
    """HBM bytes per launch of `kernel` over the timed window from the calibrated FETCH_SIZE /
    WRITE_SIZE runs kept in profiles/ -- only when those runs counted this exact window (same
    workload, burn-in, warm-up and steps); otherwise None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None
    if prof.get("regime") != tag:
        return None
    return prof.get("kernels", {}).get(kernel, {}).get("hbm_bytes_per_launch")


# ---- CPU baseline (the reference's CPU step on the host's cores) ------------------------------

def _port_sample(cfg, csr, steps, conductance=False, state=None, budget_s=12.0):
    """Time `steps` steps (fewer when one step exceeds the budget) of the CPU port on a CSR
    network; returns seconds per step, steps run and the port's own spikes and events."""
    from oracle import cpu_port
    rp, col, w, d = csr
    net = cpu_port.PortNet(cfg, rp, col, w, d, conductance=conductance)
    if state is not None:
        net.set_state(*state)
    s1, sp, ev = net.run(1)
    done = 1
    k = int(max(0, min(steps - 1, (budget_s - s1) / max(s1, 1e-9))))
    if k:
        s2, sp2, ev2 = net.run(k)
        s1, sp, ev, done = s1 + s2, sp + sp2, ev + ev2, done + k
    net.close()
    return s1 / done, done, sp, ev


def _cpu_entry(value, unit, sample, extra=None):
    from oracle import cpu_port
    info = cpu_port.host_info()
    out = {"value": value, "unit": unit, "cores": info["threads"], "kind": "port",
           "sample": sample, "cpu_model": info["cpu_model"], "nproc": info["nproc"]}
    if extra:
        out.update(extra)
    return out


# ---- workloads -------------------------------------------------------------------------------

def bench_headline(ns, args):
    """The 1M/100M Izhikevich + pair STDP network: burn-in, warm-up, K timed bench steps (graph
    replay), then the same steps replayed with per-kernel events."""
    spec = ns.configs.izh_stdp_network(n=args.neurons)
    dt = spec.cfg.dt_ms
    per = int(round(args.step_ms / dt))  # simulation steps per bench step
    t0 = time.perf_counter()
    eng = ns.SimulationEngine.from_spec(spec)
    build_s = time.perf_counter() - t0
    log(f"bench: {spec.name} n={spec.n} m={spec.m} built on device in {build_s:.2f}s")
    if args.split:
        eng.set_prop_split(args.split)
    set_headline_delivery(ns, eng, args.delivery)
    G = args.graph
    eng.set_graph_steps(G)
    eng.step(int(round(args.burnin_ms / dt)))
    tuned = None
    if args.delivery == "auto":
        eng.autotune(200)
        r = eng.last_tune
        tuned = {"prop_split": int(r.prop_split), "delivery": ("atomic", "binned", "pull")[
            int(eng.delivery)], "tile_entries": int(r.tile_entries), "row_split": int(r.row_split),
            "syn_block": int(r.syn_block), "update_width": int(r.update_width)}
        eng.set_graph_steps(G)
    eng.step(args.warmup * per)
    log(f"bench: burn-in {args.burnin_ms:g} ms + warmup {args.warmup} x {args.step_ms:g} ms done")
    K = args.steps
    S = K * per  # simulation steps timed
    Gt = min(G, S) if G else 0
    eng.set_graph_steps(Gt)
    eng.snapshot_save()
    if Gt:
        eng.step(Gt)  # capture the window's graph outside the timed region
        eng.snapshot_restore()
    t0 = time.perf_counter()
    st = eng.step(S)
    wall = time.perf_counter() - t0
    log(f"bench: timed {K} x {args.step_ms:g} ms in {wall:.4f}s (device {st.wall_ms / 1e3:.4f}s)")
    sim_s = S * dt / 1000.0
    out = {
        "metric": "syn events/s & x-realtime @1M/100M Izh+STDP; HH surrogate vs plain speedup",
        "value": st.events / wall,
        "unit": "syn events/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox-generated network and drive)",
        "config": {
            "workload": "izhikevich_pair_stdp_1M_100M",
            "neurons": spec.n, "synapses": spec.m, "dt_ms": dt,
            "v_substeps": spec.cfg.v_substeps, "delays_ms": [1, 20],
            "bench_step": f"simulate_ms({args.step_ms:g}): {per} simulation steps of {dt:g} ms",
            "burnin_ms": args.burnin_ms,
            "t_start_ms": args.burnin_ms + args.warmup * per * dt,
            "sim_ms": sim_s * 1e3, "accumulate": "int64 fixed point 2^-32",
            "delivery": args.delivery if tuned is None else {"auto": tuned},
            "launch": ({"graph_steps": Gt, "graph_launches": S // Gt, "direct_steps": S % Gt}
                       if Gt else {"graph_steps": 0, "graph_launches": 0, "direct_steps": S}),
        },
        "x_realtime": sim_s / wall,
        "ms_per_sim_step": wall * 1e3 / S,
        "spikes": st.spikes,
        "events": st.events,
        "events_per_sim_step": st.events / S,
        "mean_rate_hz": st.spikes / spec.n / sim_s,
        "build_s": build_s,
    }
    if not args.no_profile:
        # the same steps again, from the snapshot, as direct launches with per-kernel events
        eng.snapshot_restore()
        eng.set_graph_steps(0)
        eng.set_profiling(True)
        pst = eng.step(S)
        eng.set_profiling(False)
        assert pst.events == st.events and pst.spikes == st.spikes, "replay differs"
        binned = int(eng.delivery) == ns.abi.DELIVERY_BINNED
        syn = "k_synapses_bin" if binned else "k_synapses"
        kinds = {0: "k_update", 1: syn}
        if binned:
            kinds.update({2: "k_bin_scan+k_bin_place", 3: "k_bin_arrive"})
        times = {k: eng.kernel_time(k) for k in kinds}
        ltp_slots, ltd_slots = times[1][2], eng.kernel_time(2)[2]
        inv = 1.0 / S
        ev = st.events * inv
        upd_b = neuron_bytes(spec.cfg, spec.n, st.spikes * inv)
        # the 8 B per event of the synapse pass is its ring atomic (atomic delivery) or its
        # record store (binned)
        prop_b, syn_b = synapse_bytes(ev, st.spikes * inv, ltp_slots * inv, ltd_slots * inv)
        kt = {kinds[k]: {"avg_us": (ms / max(n, 1)) * 1e3, "launches": n}
              for k, (ms, n, _) in times.items()}
        kt["k_update"]["alg_bytes_per_launch"] = upd_b
        kt[syn]["alg_bytes_per_launch"] = syn_b
        kt[syn]["propagation_bytes_per_launch"] = prop_b
        kt[syn]["events_per_launch"] = ev
        kt[syn]["ltp_slots_per_launch"] = ltp_slots * inv
        kt[syn]["ltd_slots_per_launch"] = ltd_slots * inv
        if binned:
            # placement reads and writes each 8 B record; the arrival reads it (a lower bound:
            # the ring entries it adds into are not counted)
            kt["k_bin_scan+k_bin_place"]["alg_bytes_per_launch"] = 16.0 * ev
            kt["k_bin_arrive"]["alg_bytes_per_launch"] = 8.0 * ev
        for k in kinds.values():
            kt[k]["achieved_GBs"] = (kt[k]["alg_bytes_per_launch"]
                                     / (max(kt[k]["avg_us"], 1e-9) * 1e-6) / 1e9)
        if not binned:
            # one 64-bit ring atomic per event: rate against the device's measured random-atomic
            # throughput (profiles/atomic_ceiling.md)
            kt[syn]["atomics_G_per_s"] = ev / (kt[syn]["avg_us"] * 1e3)
            kt[syn]["atomic_ceiling_G_per_s"] = 23.0
        dom = max(kinds.values(), key=lambda k: kt[k]["avg_us"])
        out["roofline"] = {
            "kernel": dom, "bound": "hbm", "achieved": kt[dom]["achieved_GBs"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kt[dom]["achieved_GBs"] / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom, regime_tag(args)),
        }
        out["kernels"] = kt
        kern_ms = sum(kt[k]["avg_us"] for k in kinds.values()) / 1e3
        out["profiled_segment"] = {"sim_steps": S, "launch": "direct, hipEvents per kernel",
                                   "same_steps_as_timed": True,
                                   "kernel_ms_per_sim_step": kern_ms,
                                   "ms_per_sim_step": pst.wall_ms / S}
    cpu_in = None
    if not args.no_cpu:
        eng.snapshot_restore()
        s = eng.read_state(["v", "u"])
        cpu_in = (eng.read_csr(), (s["v"], s["u"]), spec, st.events / S)
    eng.close()
    return out, cpu_in


def cpu_headline(cpu_in, budget_s):
    """The CPU port on the same network (weights after the burn-in) from the device's state at
    the start of the timed window; converted to the metric's unit with the device's events per
    simulation step over the timed window (the port's O(M) passes make its step cost nearly
    independent of activity)."""
    csr, state, spec, ev_per_step = cpu_in
    log("cpu baseline: 1M/100M port")
    s_step, done, sp, ev = _port_sample(spec.cfg, csr, 8, state=state, budget_s=budget_s)
    return _cpu_entry(
        ev_per_step / s_step, "syn events/s",
        f"{done} simulation steps (1 ms) of the 1M/100M network in the reference's CPU step "
        f"(per-synapse AoS, per-synapse STDP traces, O(M) passes; OpenMP neuron phase) from the "
        f"device's state at the start of the timed window: {s_step * 1e3:.0f} ms per step; "
        f"value = the device's events per step over the timed window / CPU seconds per step",
        {"ms_per_sim_step": s_step * 1e3, "x_realtime": 1e-3 / s_step,
         "port_events_per_step": ev / done})


def bench_lif10k(ns, cpu=True, budget_s=8.0):
    """The CPU-runnable 10k LIF network: 1 s simulated (10^4 steps of 0.1 ms) at the stated
    drive (silent) and with the drive x60 (active)."""
    from neuralsim_amd import configs
    res = {}
    for key, scale in (("stated_drive", 1.0), ("drive_x60", 60.0)):
        spec = configs.lif_network(n=10_000, drive_scale=scale)
        with ns.SimulationEngine.from_spec(spec) as eng:
            eng.set_graph_steps(100)
            eng.step(100)  # captures the 100-step graph
            t0 = time.perf_counter()
            st = eng.step(10_000 - 100)
            wall = time.perf_counter() - t0
            r = {"steps": 10_000 - 100, "ms_per_step": wall * 1e3 / (10_000 - 100),
                 "x_realtime": (10_000 - 100) * 0.1e-3 / wall, "syn_events_per_s": st.events / wall,
                 "rate_hz": st.spikes / spec.n / ((10_000 - 100) * 1e-4)}
            if cpu:
                csr = eng.read_csr()
                s_step, done, sp, ev = _port_sample(spec.cfg, csr, 4000, budget_s=budget_s / 2)
                r["cpu_baseline"] = _cpu_entry(
                    1e-4 / s_step, "x realtime",
                    f"{done} steps (0.1 ms) of the same network in the reference's CPU step",
                    {"ms_per_step": s_step * 1e3, "port_events_per_step": ev / done})
            res[key] = r
    return res


def _timed(eng, steps, graph, epoch=128):
    eng.set_epoch(epoch)
    eng.set_graph_steps(graph)
    eng.step(max(graph, 1) * 2)  # capture and warm up
    t0 = time.perf_counter()
    st = eng.step(steps)
    return time.perf_counter() - t0, st


def _kernel_us(eng, steps):
    """Per-step device time of the update and synapse kernels over `steps` direct-launch steps
    (an epoch's launch covers several steps)."""
    eng.set_graph_steps(0)
    eng.set_profiling(True)
    eng.step(steps)
    eng.set_profiling(False)
    ms_u, _, _ = eng.kernel_time(0)
    ms_s, _, _ = eng.kernel_time(1)
    return ms_u * 1e3 / steps, ms_s * 1e3 / steps


HH_LEGS = ("surrogate", "plain_dt0.01", "plain_dt0.05_x5", "accuracy")


def bench_hh(ns, n=200_000, sim_ms=100.0, graph=100, legs=HH_LEGS, cpu=False, budget_s=8.0):
    """HH network of 200k neurons: the learned surrogate (dt 0.05 ms) against the plain
    integrator at dt 0.01 ms and against plain five-sub-step fusion at dt 0.05 ms, each run
    with one launch pair per step ("per_step") and in epochs of d_min steps ("epoch": 1 ms of
    simulated time per update launch, the shortest delay); per-kernel MFMA rate of the surrogate
    update; spike displacement of the surrogate against the plain FP32 integrator on an
    uncoupled population under the same drive.  `legs` selects a subset (one engine per process
    keeps a kernel trace to that leg)."""
    from neuralsim_amd import configs, surrogate
    from neuralsim_amd.metrics import spike_displacement
    res = {}
    params = norm = None

    def both(eng, name, dt, steps):
        out = {}
        for mode, ep in (("per_step", 1), ("epoch", 128)):
            wall, st = _timed(eng, steps, graph, ep)
            uk, sk = _kernel_us(eng, min(steps, 2000))
            out[mode] = {"epoch_steps": eng.epoch, "ms_per_sim_s": wall / sim_ms * 1e6,
                         "us_per_step": wall / steps * 1e6, "update_us_per_step": uk,
                         "synapse_us_per_step": sk,
                         "rate_hz": st.spikes / n / (sim_ms / 1e3)}
        return out

    if "surrogate" in legs:
        spec = configs.hh_surrogate_network(n=n)
        with ns.SimulationEngine.from_spec(spec) as eng:
            fit = surrogate.fit(eng)
            params, norm = eng.surrogate_params()
            res["fit"] = {"samples": fit["samples"], "generate_s": fit["generate_s"],
                          "train_s": fit["train_s"], "steps": len(fit["loss"]),
                          "final_loss": float(fit["loss"][-200:].mean())}
            res["surrogate"] = {"dt_ms": 0.05, **both(eng, "surrogate", 0.05,
                                                      int(round(sim_ms / 0.05)))}
            res["surrogate"]["fallback_neuron_steps"] = eng.surrogate_fallbacks()
            us = res["surrogate"]["epoch"]["update_us_per_step"]
            flop = 2.0 * (5 * 64 + 64 * 64 + 2 * 64 * 4) * n  # W3 runs as two bf16 words
            res["surrogate"]["mlp_tflops"] = flop / (us * 1e-6) / 1e12
            res["surrogate"]["mlp_frac_of_bf16_peak"] = flop / (us * 1e-6) / BF16_PEAK
    for name, dt, sub in [("plain_dt0.01", 0.01, 1), ("plain_dt0.05_x5", 0.05, 5)]:
        if name not in legs:
            continue
        with ns.SimulationEngine.from_spec(configs.hh_network(n=n, dt_ms=dt,
                                                              hh_substeps=sub)) as eng:
            res[name] = {"dt_ms": dt, "substeps": sub,
                         **both(eng, name, dt, int(round(sim_ms / dt)))}
            if cpu and name == "plain_dt0.01":
                log("cpu baseline: HH 200k/20M port")
                csr = eng.read_csr()
                s_step, done, sp, ev = _port_sample(configs.hh_network(n=n).cfg, csr, 40,
                                                    conductance=True, budget_s=budget_s)
                res["cpu_baseline"] = _cpu_entry(
                    1e-5 / s_step, "x realtime",
                    f"{done} steps (0.01 ms) of the 200k/20M network in the reference's CPU step in "
                    f"conductance mode (per-synapse conductances decayed and summed every step, "
                    f"OpenMP neuron phase); value = 0.01 ms / seconds per step",
                    {"ms_per_step": s_step * 1e3, "ms_per_sim_s": s_step * 1e5 * 1e3})
    if "surrogate" in res:
        sur = res["surrogate"]
        for name in ("plain_dt0.01", "plain_dt0.05_x5"):
            if name in res:
                # same launch structure on both sides, and each side per step
                res[f"speedup_vs_{name}"] = (res[name]["epoch"]["ms_per_sim_s"]
                                             / sur["epoch"]["ms_per_sim_s"])
                res[f"speedup_vs_{name}_per_step"] = (res[name]["per_step"]["ms_per_sim_s"]
                                                      / sur["per_step"]["ms_per_sim_s"])
        if "plain_dt0.01" in res:
            res["speedup_epoch_vs_plain_dt0.01_per_step"] = (
                res["plain_dt0.01"]["per_step"]["ms_per_sim_s"] / sur["epoch"]["ms_per_sim_s"])
    if "accuracy" in legs and params is not None:
        # 2000 uncoupled neurons, 500 ms, same OU drive
        kw = dict(k_in=1, w_exc=(0.0, 0.0), w_inh=(0.0, 0.0), record_raster=True,
                  raster_capacity=1 << 22)
        u = configs.hh_surrogate_network(n=2000, **kw)
        with ns.SimulationEngine.from_spec(u) as eng:
            eng.set_surrogate_params(params, norm)
            eng.step(10_000)
            a = eng.raster()
        with ns.SimulationEngine.from_spec(configs.hh_network(n=2000, dt_ms=0.05, hh_substeps=5,
                                                              **kw)) as eng:
            eng.step(10_000)
            b = eng.raster()
        res["displaced_vs_plain_fp32"] = spike_displacement(b[0], b[1], a[0], a[1], 0.05, 0.5)
    return res


def bench_population(ns, steps=2000, warm=200, tune_steps=100, cpu=False, budget_s=12.0):
    """256 Izhikevich networks (N log-uniform in [1k, 64k], in-degree uniform in [20, 200]) as
    one batched population, 2 s at dt 1 ms: the first `warm` steps (the networks leave their
    quiet initial state), autotuning from that state (trials restore it), then the remaining
    steps timed; RateHomeostasis(5 Hz) fitness over all 2 s."""
    from neuralsim_amd import configs
    from neuralsim_amd.population import RateHomeostasis
    import numpy as np
    base = configs.population_base()
    nets = configs.population_nets(256)
    t0 = time.perf_counter()
    with ns.BatchedPopulation(base, nets) as pop:
        build_s = time.perf_counter() - t0
        pop.set_graph_steps(10)
        pop.step(warm)
        split, graph, _ = pop.autotune(tune_steps)
        r = pop.last_tune
        row_split, syn_block, width = int(r.row_split), int(r.syn_block), int(r.update_width)
        trials = [{k: (round(v, 1) if k == "us_per_step" else int(v)) for k, v in t.items()}
                  for t in pop.tune_trials()]
        pop.snapshot_save()
        pop.step(graph)  # capture the tuned window's graph outside the timed region
        pop.snapshot_restore()
        t0 = time.perf_counter()
        st = pop.step(steps - warm)
        wall = time.perf_counter() - t0
        spikes, _ = pop.net_stats()
        deliv, tile = pop.delivery, int(pop.last_tune.tile_entries)
        # the timed steps again from the snapshot as direct launches bracketed by hipEvents
        # (the replay repeats them bit for bit).  Only the sum over a step's kernels is kept:
        # between adjacent kernels the event stamps shift time from one to the other (a kernel
        # trace of the same launches: population_256_izh_kernels.md)
        pop.snapshot_restore()
        pop.set_profiling(True)
        pst = pop.step(steps - warm)
        pop.set_profiling(False)
        assert pst.events == st.events and pst.spikes == st.spikes
        kt = sum(pop.kernel_time(k)[0] for k in range(4)) * 1e3 / (steps - warm)
    n = np.array([b.n_neurons for b in nets], np.float64)
    m = np.array([b.n_neurons * b.k_in for b in nets], np.float64)
    rate = spikes / n / (steps * base.dt_ms / 1e3)
    fit = RateHomeostasis(5.0).fitness(rate)
    res = {"networks": len(nets), "neurons": int(n.sum()), "synapses": int(m.sum()),
           "steps": steps, "build_s": build_s,
           "tuned": {"prop_split": split, "graph_steps": graph,
                     "delivery": ("atomic", "binned", "pull")[deliv], "tile_entries": tile,
                     "row_split": row_split, "syn_block": syn_block, "update_width": width,
                     "persistent": int(r.persistent), "trials": trials},
           "timed_steps": steps - warm,
           "replay_kernel_us_per_step": kt,
           "ms_per_step": wall * 1e3 / (steps - warm),
           "x_realtime": (steps - warm) * base.dt_ms / 1e3 / wall,
           "syn_events_per_s": st.events / wall, "spikes": int(spikes.sum()),
           "rate_hz": {"min": float(rate.min()), "median": float(np.median(rate)),
                       "max": float(rate.max())},
           "fitness_mean": float(fit.mean()), "fitness_best": float(fit.max())}
    if cpu:
        res["cpu_baseline"] = cpu_population(base, nets, m, budget_s)
    return res


def bench_population_direct(ns, delivery="binned", warm=200, steps=200, split=8, tile=16384):
    """The population with a fixed shape (by default the one the autotuner picks on this
    machine: binned delivery with 16384-entry tiles, propagation split 8; or pull or atomic
    delivery), no tuning and no graphs: every step is direct launches, so a kernel trace sees
    each one.  `warm` steps, then `steps` timed."""
    from neuralsim_amd import abi, configs
    mode = {"atomic": abi.DELIVERY_ATOMIC, "binned": abi.DELIVERY_BINNED,
            "pull": abi.DELIVERY_PULL}[delivery]
    with ns.BatchedPopulation(configs.population_base(), configs.population_nets(256)) as pop:
        pop.set_graph_steps(0)
        pop.set_prop_split(split)
        pop.set_delivery(mode, 0, tile if mode == abi.DELIVERY_BINNED else 0)
        pop.step(warm)
        t0 = time.perf_counter()
        st = pop.step(steps)
        wall = time.perf_counter() - t0
    return {"delivery": delivery, "split": split, "warm_steps": warm, "steps": steps,
            "ms_per_step": wall * 1e3 / steps, "events_per_step": st.events / steps}


def cpu_population(base, nets, m, budget_s, stride=16, steps=20):
    """Population evaluation on the CPU port (one engine per genome, genomes in parallel as
    population.rs:116-124 does with rayon) over every `stride`-th network; the port's cost is
    O(M) per step, so the whole population's step is extrapolated by synapse count, assuming
    the genomes spread over all cores."""
    import numpy as np
    from neuralsim_amd import configs
    from oracle import cpu_port, oracle
    log("cpu baseline: population sample")
    idx = list(range(0, len(nets), stride))
    ports = []
    for i in idx:
        cfg = configs.population_single(base, nets[i])
        spec = configs.NetSpec("pop", cfg, nets[i].k_in, (nets[i].w_exc_lo, nets[i].w_exc_hi),
                               (nets[i].w_inh_lo, nets[i].w_inh_hi), (nets[i].d_lo, nets[i].d_hi), 1)
        src, dst, w, d = oracle.generate_list(spec)
        rp, col, perm = oracle.finalize(spec.n, src, dst)
        ports.append(cpu_port.PortNet(cfg, rp, col, w[perm], d[perm]))
    wall, secs = cpu_port.run_many(ports, steps)
    for p in ports:
        p.close()
    threads = cpu_port.host_info()["threads"]
    per_syn_step = secs.sum() / (m[idx].sum() * steps)  # one thread's seconds per synapse-step
    pop_step = per_syn_step * m.sum() / threads
    return _cpu_entry(
        1e-3 / pop_step, "x realtime",
        f"{len(idx)} of the 256 networks (every {stride}th), {steps} steps each, genomes in "
        f"parallel on {threads} threads; population step = per-synapse cost x all synapses / "
        f"threads = {pop_step * 1e3:.0f} ms (2 s simulated: {pop_step * 2000:.0f} s)",
        {"ms_per_step": pop_step * 1e3, "sample_wall_s": wall})


def bench_large(ns, n=100_000_000, steps=300, drive_scale=1.0):
    """10^8 LIF neurons, 10^10 synapses, pair STDP on E->E: device build, then steps.  At the
    workload's drive the network is silent and the neuron sweep (HBM-bound) is the whole step;
    drive_scale multiplies the Poisson weight to run the same network active.  The timed steps
    are direct launches bracketed by per-kernel hipEvents, so the kernel times are those of the
    timed steps (at 0.4-3 ms per step a graph saves nothing measurable, and a snapshot of this
    network would not fit beside it in HBM)."""
    from neuralsim_amd import configs
    spec = configs.lif_large_network(n=n, drive_scale=drive_scale)
    t0 = time.perf_counter()
    with ns.SimulationEngine.from_spec(spec) as eng:
        build_s = time.perf_counter() - t0
        eng.set_graph_steps(0)
        eng.step(50)
        eng.set_profiling(True)
        t0 = time.perf_counter()
        st = eng.step(steps)
        wall = time.perf_counter() - t0
        eng.set_profiling(False)
        ms_u, n_u, _ = eng.kernel_time(0)
        ms_s, n_s, _ = eng.kernel_time(1)
    upd_us = ms_u / max(n_u, 1) * 1e3
    # a silent segment (no spike in it nor in the D - 1 steps before) skips the ring loads
    upd_b = neuron_bytes(spec.cfg, spec.n, st.spikes / steps, ring=st.spikes > 0)
    sim_s = steps * spec.cfg.dt_ms / 1e3
    return {"neurons": spec.n, "synapses": spec.m, "drive_scale": drive_scale,
            "build_s": build_s, "steps": steps, "launch": "direct, hipEvents per kernel",
            "ms_per_step": wall * 1e3 / steps, "x_realtime": sim_s / wall,
            "syn_events_per_s": st.events / wall, "spikes": int(st.spikes),
            "k_update": {"avg_us": upd_us, "alg_bytes_per_launch": upd_b,
                         "achieved_GBs": upd_b / (upd_us * 1e-6) / 1e9,
                         "frac_of_hbm_peak": upd_b / (upd_us * 1e-6) / 1e9 / HBM_PEAK_GBS},
            "k_synapses_avg_us": ms_s / max(n_s, 1) * 1e3,
            "kernel_ms_per_step": (ms_u + ms_s) / steps}


def cpu_large(ns, budget_s):
    """The 10^8-neuron network is infeasible on the CPU port (its AoS synapse state alone is
    about 1 TB): the same network at 1/100 scale (10^6 LIF neurons, 10^8 synapses, pair STDP
    E->E), timed over a step window; reported as such, not extrapolated across scale."""
    from neuralsim_amd import configs
    spec = configs.lif_large_network(n=1_000_000, drive_scale=60.0)
    with ns.SimulationEngine.from_spec(spec) as eng:
        csr = eng.read_csr()
    log("cpu baseline: LIF 10^6 / 10^8 port (1/100 scale)")
    s_step, done, sp, ev = _port_sample(spec.cfg, csr, 8, budget_s=budget_s)
    return _cpu_entry(
        1e-4 / s_step, "x realtime",
        f"1/100 scale: {done} steps (0.1 ms) of a 10^6-neuron / 10^8-synapse LIF + pair STDP "
        f"network (drive x60) in the reference's CPU step; not extrapolated to 10^8 neurons",
        {"ms_per_step": s_step * 1e3, "scale": 0.01})


def probe_atomics(reps=20):
    """Device-wide throughput of random integer atomic adds: 2^24 adds per launch to uniformly
    random entries of a 168 MB int64 array (the 1M network's ring: 21 slots x 10^6 neurons) and
    of an 84 MB int32 array, through torch's scatter_add_ (one hardware atomic per element; its
    index and value streams are coalesced).  Returns G atomics per second."""
    import torch
    dev = torch.device("cuda:0")
    n_adds, out = 1 << 24, {}
    gen = torch.Generator(device=dev).manual_seed(7)
    for name, dtype in (("int64", torch.int64), ("int32", torch.int32)):
        tgt = torch.zeros(21 * 1_000_000, dtype=dtype, device=dev)
        idx = torch.randint(0, tgt.numel(), (n_adds,), generator=gen, device=dev)
        val = torch.ones(n_adds, dtype=dtype, device=dev)
        tgt.scatter_add_(0, idx, val)  # warm up
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            tgt.scatter_add_(0, idx, val)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        assert int(tgt.sum().item()) == (reps + 1) * n_adds
        out[name] = {"adds_per_launch": n_adds, "target_bytes": tgt.numel() * tgt.element_size(),
                     "us_per_launch": ms * 1e3, "G_atomics_per_s": n_adds / (ms * 1e-3) / 1e9}
    return out


LEGS = ("headline", "lif10k", "hh_surrogate", "hh_plain", "hh_plain_x5", "population",
        "population_direct", "large", "large_active")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100,
                    help="timed bench steps (each one simulate_ms(--step-ms))")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--step-ms", type=float, default=100.0,
                    help="simulated ms per bench step of the headline network")
    ap.add_argument("--burnin-ms", type=float, default=1000.0,
                    help="simulated ms run before the warm-up (out of the quiet initial state)")
    ap.add_argument("--neurons", type=int, default=1_000_000)
    ap.add_argument("--graph", type=int, default=100, help="steps per captured hipGraph (0: off)")
    ap.add_argument("--no-profile", action="store_true", help="no per-kernel replay")
    ap.add_argument("--split", type=int, default=0, help="wavefronts per neuron block (0: default)")
    ap.add_argument("--delivery", choices=("binned", "atomic", "auto"), default="binned",
                    help="spike delivery of the headline network (auto: the autotuner's choice)")
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of CPU-port stepping per workload sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hh", action="store_true", help="skip the HH surrogate comparison")
    ap.add_argument("--no-pop", action="store_true", help="skip the batched population")
    ap.add_argument("--no-large", action="store_true", help="skip the 10^8-neuron network")
    ap.add_argument("--no-lif10k", action="store_true", help="skip the 10k LIF network")
    ap.add_argument("--leg", choices=LEGS,
                    help="run only this workload (one engine per process, for kernel traces)")
    ap.add_argument("--pop-delivery", choices=("binned", "pull", "atomic"), default="binned",
                    help="delivery of the population_direct leg")
    ap.add_argument("--probe-atomics", action="store_true",
                    help="measure the device's random int64 / int32 atomic-add throughput only")
    ap.add_argument("--counters-window", action="store_true",
                    help="headline network only: burn-in, warm-up and the timed steps as direct "
                         "launches, nothing else (for rocprofv3 --pmc over that window)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU (asked for --gpus {args.gpus}); runs on several "
              "GPUs are not supported", file=sys.stderr)
        sys.exit(2)

    if args.probe_atomics:
        print(json.dumps({"atomics": probe_atomics()}), flush=True)
        return
    import neuralsim_amd as ns
    if args.counters_window:
        spec = ns.configs.izh_stdp_network(n=args.neurons)
        per = int(round(args.step_ms / spec.cfg.dt_ms))
        before = int(round(args.burnin_ms / spec.cfg.dt_ms)) + args.warmup * per
        with ns.SimulationEngine.from_spec(spec) as eng:
            set_headline_delivery(ns, eng, args.delivery)
            eng.set_graph_steps(0)
            eng.step(before)
            st = eng.step(args.steps * per)
        print(json.dumps({"counters_window": regime_tag(args),
                          "k_synapses_launches_before": before,
                          "k_synapses_launches_window": args.steps * per,
                          "events": st.events, "spikes": st.spikes}), flush=True)
        return
    if args.leg and args.leg != "headline":
        res = {"leg": args.leg}
        if args.leg == "lif10k":
            res.update(bench_lif10k(ns, cpu=False))
        elif args.leg.startswith("hh"):
            leg = {"hh_surrogate": "surrogate", "hh_plain": "plain_dt0.01",
                   "hh_plain_x5": "plain_dt0.05_x5"}[args.leg]
            res.update(bench_hh(ns, legs=(leg,)))
        elif args.leg == "population":
            res.update(bench_population(ns))
        elif args.leg == "population_direct":
            res.update(bench_population_direct(ns, args.pop_delivery))
        else:
            scale, steps = (1.0, 300) if args.leg == "large" else (60.0, 100)
            res.update(bench_large(ns, steps=steps, drive_scale=scale))
        print(json.dumps(res), flush=True)
        return

    out, cpu_in = bench_headline(ns, args)
    only = args.leg == "headline"
    cpu = not args.no_cpu and not only
    if not args.no_lif10k and not only:
        log("bench: 10k LIF network, 1 s")
        out["lif_10k"] = bench_lif10k(ns, cpu=cpu, budget_s=args.cpu_budget)
    if not args.no_hh and not only:
        log("bench: HH surrogate vs plain (200k neurons)")
        out["hh_surrogate"] = bench_hh(ns, cpu=cpu, budget_s=args.cpu_budget)
        log(f"bench: HH speedup vs plain dt 0.01 ms "
            f"{out['hh_surrogate'].get('speedup_vs_plain_dt0.01', 0):.2f}")
    if not args.no_pop and not only:
        log("bench: batched population (256 networks)")
        out["population"] = bench_population(ns, cpu=cpu, budget_s=args.cpu_budget)
        log(f"bench: population {out['population']['ms_per_step']:.3f} ms/step")
    if not args.no_large and not only:
        log("bench: 10^8 LIF neurons, 10^10 synapses")
        for key, scale, steps in [("large_scale", 1.0, 300), ("large_scale_active", 60.0, 100)]:
            try:
                out[key] = bench_large(ns, steps=steps, drive_scale=scale)
                log(f"bench: {key} {out[key]['ms_per_step']:.3f} ms/step")
            except ns.NsxError as e:  # recorded, not fatal: the headline line still prints
                out[key] = {"error": str(e)}
                log(f"bench: {key} failed: {e}")
        if cpu and "error" not in out.get("large_scale", {}):
            out["large_scale"]["cpu_baseline"] = cpu_large(ns, args.cpu_budget)
    if cpu_in is not None and not args.no_cpu:
        out["cpu_baseline"] = cpu_headline(cpu_in, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
