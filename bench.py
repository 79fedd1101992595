#!/usr/bin/env python3
"""Benchmark: syn events/s and x-realtime of the 1M-neuron / 100M-synapse Izhikevich + pair-STDP
network on one MI355X, with the propagation/update kernel roofline and a CPU baseline.

A step is one 1 ms simulation step of the whole network (neuron update, propagation with delays,
STDP).  `value` = synaptic events (sum of out-degrees of spiking rows) per wall second over the
K timed steps, inputs (network, state) resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


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


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the calibrated FETCH_SIZE / WRITE_SIZE runs kept in
    profiles/ (rocprofv3 --pmc; commands in that file), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_izh_stdp_1M_100M.json")
    try:
        with open(path) as f:
            return json.load(f)["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(spec, gpu_events_per_step, steps):
    """FP64 restatement of the reference's step (per-synapse STDP traces, serial O(M) synapse
    passes as engine.rs:397-537) timed on one host core over a bounded window of the same
    network; converted to the metric's unit with the device run's mean events per step."""
    from oracle import oracle
    t0 = time.perf_counter()
    src, dst, w, d = oracle.generate_list(spec)
    rp, col, perm = oracle.finalize(spec.n, src, dst)
    w, d = w[perm], d[perm]
    del src, dst, perm
    ref = oracle.RefNet(spec.cfg, rp, col, w, d)
    del rp, col, w, d
    setup = time.perf_counter() - t0
    log(f"cpu baseline: network ready in {setup:.1f}s, timing {steps} steps")
    t0 = time.perf_counter()
    ref.run(steps, cap=1 << 24)
    dt = time.perf_counter() - t0
    ref.close()
    s_per_step = dt / steps
    return {
        "value": gpu_events_per_step / s_per_step,
        "unit": "syn events/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{steps} steps (from t=0) of the same 1M/100M network in the oracle's FP64 "
                   f"restatement of SimulationEngine::step with per-synapse STDP traces, "
                   f"{s_per_step * 1e3:.0f} ms/step on 1 core; events/s = device mean events "
                   f"per step / CPU seconds per step"),
        "ms_per_step": s_per_step * 1e3,
    }


def _timed(eng, steps, graph):
    eng.set_graph_steps(graph)
    eng.step(max(graph, 1) * 2)  # capture and warm up
    t0 = time.perf_counter()
    st = eng.step(steps)
    return time.perf_counter() - t0, st


HH_LEGS = ("surrogate", "plain_dt0.01", "plain_dt0.05_x5", "accuracy")


def bench_hh(ns, n=200_000, sim_ms=100.0, graph=50, legs=HH_LEGS):
    """HH network of 200k neurons: the learned surrogate (dt 0.05 ms) against the plain
    integrator at dt 0.01 ms and against plain five-sub-step fusion at dt 0.05 ms; per-kernel
    MFMA rate of the surrogate update; spike displacement of the surrogate against the plain
    FP32 integrator on an uncoupled population under the same drive.  `legs` selects a subset
    (one engine per process keeps a kernel trace to that leg)."""
    from neuralsim_amd import configs, surrogate
    from neuralsim_amd.metrics import spike_displacement
    res = {}
    if "surrogate" not in legs:
        for name, dt, sub in [("plain_dt0.01", 0.01, 1), ("plain_dt0.05_x5", 0.05, 5)]:
            if name in legs:
                with ns.SimulationEngine.from_spec(configs.hh_network(n=n, dt_ms=dt,
                                                                      hh_substeps=sub)) as eng:
                    steps = int(round(sim_ms / dt))
                    wall, st = _timed(eng, steps, graph)
                    res[name] = {"dt_ms": dt, "substeps": sub,
                                 "ms_per_sim_s": wall / sim_ms * 1e6,
                                 "us_per_step": wall / steps * 1e6,
                                 "rate_hz": st.spikes / n / (sim_ms / 1e3)}
        return res
    spec = configs.hh_surrogate_network(n=n)
    with ns.SimulationEngine.from_spec(spec) as eng:
        fit = surrogate.fit(eng)
        params, norm = eng.surrogate_params()
        res["fit"] = {"samples": fit["samples"], "generate_s": fit["generate_s"],
                      "train_s": fit["train_s"], "steps": len(fit["loss"]),
                      "final_loss": float(fit["loss"][-200:].mean())}
        steps = int(round(sim_ms / 0.05))
        wall, st = _timed(eng, steps, graph)
        res["surrogate"] = {"dt_ms": 0.05, "ms_per_sim_s": wall / sim_ms * 1e6,
                            "us_per_step": wall / steps * 1e6,
                            "rate_hz": st.spikes / n / (sim_ms / 1e3),
                            "fallback_neuron_steps": eng.surrogate_fallbacks()}
        eng.set_profiling(True)
        eng.step(200)
        ms, launches, _ = eng.kernel_time(0)
        eng.set_profiling(False)
        us = ms / max(launches, 1) * 1e3
        flop = 2.0 * (5 * 64 + 64 * 64 + 2 * 64 * 4) * n  # W3 runs as two bf16 words
        res["surrogate"]["update_kernel_us"] = us
        res["surrogate"]["mlp_tflops"] = flop / (us * 1e-6) / 1e12
        res["surrogate"]["mlp_frac_of_bf16_peak"] = flop / (us * 1e-6) / 2.5e15
    if tuple(legs) == ("surrogate",):
        return res
    for name, dt, sub in [("plain_dt0.01", 0.01, 1), ("plain_dt0.05_x5", 0.05, 5)]:
        with ns.SimulationEngine.from_spec(configs.hh_network(n=n, dt_ms=dt,
                                                              hh_substeps=sub)) as eng:
            steps = int(round(sim_ms / dt))
            wall, st = _timed(eng, steps, graph)
            res[name] = {"dt_ms": dt, "substeps": sub, "ms_per_sim_s": wall / sim_ms * 1e6,
                         "us_per_step": wall / steps * 1e6,
                         "rate_hz": st.spikes / n / (sim_ms / 1e3)}
    s_ms = res["surrogate"]["ms_per_sim_s"]
    res["speedup_vs_plain_dt0.01"] = res["plain_dt0.01"]["ms_per_sim_s"] / s_ms
    res["speedup_vs_plain_dt0.05_x5"] = res["plain_dt0.05_x5"]["ms_per_sim_s"] / s_ms
    # accuracy: 2000 uncoupled neurons, 500 ms, same OU drive
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


def bench_population(ns, steps=2000, warm=200, tune_steps=100):
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
        split, graph, trials = pop.autotune(tune_steps)
        t0 = time.perf_counter()
        st = pop.step(steps - warm)
        wall = time.perf_counter() - t0
        spikes, _ = pop.net_stats()
        deliv, tile = pop.delivery, int(pop.last_tune.tile_entries)
    n = np.array([b.n_neurons for b in nets], np.float64)
    rate = spikes / n / (steps * base.dt_ms / 1e3)
    fit = RateHomeostasis(5.0).fitness(rate)
    return {"networks": len(nets), "neurons": int(n.sum()),
            "synapses": int(sum(b.n_neurons * b.k_in for b in nets)), "steps": steps,
            "build_s": build_s,
            "tuned": {"prop_split": split, "graph_steps": graph,
                      "delivery": "binned" if deliv else "atomic", "tile_entries": tile,
                      "trials_us_per_step": [[int(a), int(b), round(c, 1), int(d)]
                                             for a, b, c, d in trials]},
            "timed_steps": steps - warm,
            "ms_per_step": wall * 1e3 / (steps - warm),
            "x_realtime": (steps - warm) * base.dt_ms / 1e3 / wall,
            "syn_events_per_s": st.events / wall, "spikes": int(spikes.sum()),
            "rate_hz": {"min": float(rate.min()), "median": float(np.median(rate)),
                        "max": float(rate.max())},
            "fitness_mean": float(fit.mean()), "fitness_best": float(fit.max())}


def bench_large(ns, n=100_000_000, steps=300, drive_scale=1.0, graph=50):
    """10^8 LIF neurons, 10^10 synapses, pair STDP on E->E: device build, then steps.  At the
    workload's drive the network is silent and the neuron sweep (HBM-bound) is the whole step;
    drive_scale multiplies the Poisson weight to run the same network active."""
    from neuralsim_amd import configs
    spec = configs.lif_large_network(n=n, drive_scale=drive_scale)
    t0 = time.perf_counter()
    with ns.SimulationEngine.from_spec(spec) as eng:
        build_s = time.perf_counter() - t0
        eng.set_graph_steps(graph)
        eng.step(50)
        t0 = time.perf_counter()
        st = eng.step(steps)
        wall = time.perf_counter() - t0
        eng.set_graph_steps(0)
        eng.set_profiling(True)
        pst = eng.step(50)
        eng.set_profiling(False)
        ms_u, n_u, _ = eng.kernel_time(0)
        ms_s, n_s, _ = eng.kernel_time(1)
    upd_us = ms_u / max(n_u, 1) * 1e3
    # a silent segment (no spike in it nor in the D - 1 steps before: st covers those) skips
    # the ring loads
    upd_b = neuron_bytes(spec.cfg, spec.n, pst.spikes / 50.0,
                         ring=(pst.spikes + st.spikes) > 0)
    sim_s = steps * spec.cfg.dt_ms / 1e3
    return {"neurons": spec.n, "synapses": spec.m, "drive_scale": drive_scale,
            "build_s": build_s, "steps": steps,
            "ms_per_step": wall * 1e3 / steps, "x_realtime": sim_s / wall,
            "syn_events_per_s": st.events / wall, "spikes": int(st.spikes),
            "k_update": {"avg_us": upd_us, "alg_bytes_per_launch": upd_b,
                         "achieved_GBs": upd_b / (upd_us * 1e-6) / 1e9,
                         "frac_of_hbm_peak": upd_b / (upd_us * 1e-6) / 1e9 / HBM_PEAK_GBS},
            "k_synapses_avg_us": ms_s / max(n_s, 1) * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10_000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--neurons", type=int, default=1_000_000)
    ap.add_argument("--graph", type=int, default=50, help="steps per captured hipGraph (0: off)")
    ap.add_argument("--no-profile", action="store_true", help="no per-kernel event segment")
    ap.add_argument("--profile-steps", type=int, default=2000,
                    help="steps of the direct-launch segment timed per kernel")
    ap.add_argument("--split", type=int, default=0, help="wavefronts per neuron block (0: default)")
    ap.add_argument("--cpu-steps", type=int, default=30,
                    help="steps of the CPU baseline (about 0.5 s each on one core)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hh", action="store_true", help="skip the HH surrogate comparison")
    ap.add_argument("--no-pop", action="store_true", help="skip the batched population")
    ap.add_argument("--no-large", action="store_true", help="skip the 10^8-neuron network")
    ap.add_argument("--hh-leg", choices=HH_LEGS[:3],
                    help="run only this HH leg (one engine) and print its figures")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU (asked for --gpus {args.gpus}); runs on several "
              "GPUs are not supported", file=sys.stderr)
        sys.exit(2)

    import neuralsim_amd as ns
    if args.hh_leg:
        print(json.dumps({"hh_leg": args.hh_leg,
                          **bench_hh(ns, legs=(args.hh_leg,))}), flush=True)
        return

    spec = ns.configs.izh_stdp_network(n=args.neurons, steps=args.steps)
    t0 = time.perf_counter()
    eng = ns.SimulationEngine.from_spec(spec)
    build_s = time.perf_counter() - t0
    log(f"bench: {spec.name} n={spec.n} m={spec.m} built on device in {build_s:.2f}s")
    if args.split:
        eng.set_prop_split(args.split)
    # the timed region replays captured graphs of --graph steps (the production path); a
    # separate segment of direct launches bracketed by per-kernel hipEvents on the library's
    # stream gives the kernel durations
    prof = not args.no_profile
    if args.graph:
        eng.set_graph_steps(args.graph)
    eng.step(args.warmup)
    log(f"bench: warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    st = eng.step(args.steps)
    wall = time.perf_counter() - t0
    log(f"bench: timed {args.steps} steps in {wall:.3f}s (device {st.wall_ms / 1e3:.3f}s)")
    sim_s = args.steps * spec.cfg.dt_ms / 1000.0
    events, spikes = st.events, st.spikes
    value = events / wall
    out = {
        "metric": "syn events/s & x-realtime @1M/100M Izh+STDP; HH surrogate vs plain speedup",
        "value": value,
        "unit": "syn events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox-generated network and drive)",
        "config": {
            "workload": "izhikevich_pair_stdp_1M_100M",
            "neurons": spec.n, "synapses": spec.m, "dt_ms": spec.cfg.dt_ms,
            "v_substeps": spec.cfg.v_substeps, "delays_ms": [1, 20], "sim_ms": sim_s * 1e3,
            "accumulate": "int64 fixed point 2^-32",
            "launch": f"hipGraph x{args.graph}" if args.graph else "direct",
        },
        "x_realtime": sim_s / wall,
        "spikes": spikes,
        "events": events,
        "mean_rate_hz": spikes / spec.n / sim_s,
        "build_s": build_s,
    }
    if prof:
        p_steps = min(args.steps, args.profile_steps)
        eng.set_graph_steps(0)
        eng.set_profiling(True)
        pst = eng.step(p_steps)
        eng.set_profiling(False)
        p_events, p_spikes = pst.events, pst.spikes
        kinds = {0: "k_update", 1: "k_synapses"}
        times = {k: eng.kernel_time(k) for k in kinds}
        ltp_slots, ltd_slots = times[1][2], eng.kernel_time(2)[2]
        per = 1.0 / p_steps
        upd_b = neuron_bytes(spec.cfg, spec.n, p_spikes * per)
        # propagation: 4 B col + 4 B weight + 1 B delay + 8 B int64 accumulate per event and
        # 16 B of row pointers per spiking row; potentiation: 16 B post record + 4 B weight
        # write per E->E out-slot; depression: 4 B slot + 4 B source + 8 B pre record + 8 B
        # weight read+write per E->E in-slot
        prop_b = (17.0 * p_events + 16.0 * p_spikes) * per
        syn_b = prop_b + (20.0 * ltp_slots + 24.0 * ltd_slots) * per
        kt = {kinds[k]: {"avg_us": (ms / max(n, 1)) * 1e3, "launches": n}
              for k, (ms, n, _) in times.items()}
        kt["k_update"]["alg_bytes_per_launch"] = upd_b
        kt["k_synapses"]["alg_bytes_per_launch"] = syn_b
        kt["k_synapses"]["propagation_bytes_per_launch"] = prop_b
        kt["k_synapses"]["events_per_launch"] = p_events * per
        kt["k_synapses"]["ltp_slots_per_launch"] = ltp_slots * per
        kt["k_synapses"]["ltd_slots_per_launch"] = ltd_slots * per
        for k in kinds.values():
            a = kt[k]["alg_bytes_per_launch"] / (max(kt[k]["avg_us"], 1e-9) * 1e-6) / 1e9
            kt[k]["achieved_GBs"] = a
        # the synapse pass issues one 64-bit atomic per event: rate vs the measured random
        # atomic throughput of the device (23.0 G/s, int64 add to random addresses)
        kt["k_synapses"]["atomics_G_per_s"] = p_events * per / (kt["k_synapses"]["avg_us"] * 1e3)
        kt["k_synapses"]["atomic_ceiling_G_per_s"] = 23.0
        dom = max(kinds.values(), key=lambda k: kt[k]["avg_us"])
        out["roofline"] = {
            "kernel": dom, "bound": "hbm", "achieved": kt[dom]["achieved_GBs"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kt[dom]["achieved_GBs"] / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom),
        }
        out["kernels"] = kt
        out["profiled_segment"] = {"steps": p_steps, "launch": "direct, hipEvents per kernel",
                                   "ms_per_step": pst.wall_ms / p_steps}
    if not args.no_hh:
        log("bench: HH surrogate vs plain (200k neurons)")
        out["hh_surrogate"] = bench_hh(ns)
        log(f"bench: HH speedup vs plain dt 0.01 ms {out['hh_surrogate']['speedup_vs_plain_dt0.01']:.2f}")
    if not args.no_pop:
        log("bench: batched population (256 networks)")
        out["population"] = bench_population(ns)
        log(f"bench: population {out['population']['ms_per_step']:.3f} ms/step")
    if not args.no_large:
        log("bench: 10^8 LIF neurons, 10^10 synapses")
        for key, scale, steps in [("large_scale", 1.0, 300), ("large_scale_active", 60.0, 100)]:
            try:
                out[key] = bench_large(ns, steps=steps, drive_scale=scale, graph=args.graph)
                log(f"bench: {key} {out[key]['ms_per_step']:.3f} ms/step")
            except ns.NsxError as e:  # recorded, not fatal: the headline line still prints
                out[key] = {"error": str(e)}
                log(f"bench: {key} failed: {e}")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(spec, events / args.steps, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
