"""Network specifications of the benchmark and parity workloads.

Each spec is an NsxConfig plus the fixed in-degree generator arguments.  Choices the
workload definitions leave open are fixed here (and nowhere else):

* LIF (10k and 100M neuron workloads): R = 1, so R*I is the sum of the exponentially decaying
  excitatory (tau 2 ms) and inhibitory (tau 10 ms) synaptic terms in mV; a spike of weight w
  adds w to the term of its source's type; the Poisson drive is excitatory input of its stated
  weight.  With the stated drive (50 x 10 Hz x 0.3 mV) the mean depolarisation is ~0.3 mV, far
  below the 15 mV threshold gap, so the network is essentially silent; `drive_scale` multiplies
  the drive weight for parity tests that need spikes.
* Izhikevich 1M / 100M: excitatory rows are regular spiking (0.02, 0.2, -65, 8), inhibitory rows
  fast spiking (0.1, 0.2, -65, 2); per-neuron jitter c = c0 + 15 r^2, d = d0 (1 - 0.75 r^2),
  r ~ U[0,1) (Izhikevich 2003's continuum applied to both types).  The input current of a
  1 ms step is the sum of arriving weights plus 5 N(0,1); V takes two 0.5 ms sub-steps with that
  current, u one 1 ms step from the step-start v.
* HH 200k / 20M: standard squid-axon parameters of HodgkinHuxleyNeuron::reset_state, spike =
  upward crossing of 0 mV (no artificial reset); excitatory spikes open AMPA and NMDA (Mg 1 mM,
  block evaluated at arrival), inhibitory spikes open GABA-A; delays 1-2 ms; weights chosen for
  moderate rates (the workload does not state them).
"""
import ctypes
import dataclasses

from . import abi


@dataclasses.dataclass
class NetSpec:
    name: str
    cfg: abi.NsxConfig
    k_in: int
    w_exc: tuple
    w_inh: tuple
    delay_steps: tuple
    steps: int
    description: str = ""

    def generator_args(self):
        return (self.k_in, self.w_exc[0], self.w_exc[1], self.w_inh[0], self.w_inh[1],
                self.delay_steps[0], self.delay_steps[1])

    @property
    def n(self):
        return int(self.cfg.n_neurons)

    @property
    def m(self):
        return int(self.cfg.n_neurons) * self.k_in

    @property
    def sim_ms(self):
        return self.steps * self.cfg.dt_ms


def base_config(model):
    """nsx_config_init() defaults, computed without loading the library (mirrors nsx_api.hip)."""
    c = abi.NsxConfig()
    c.struct_size = ctypes.sizeof(abi.NsxConfig)
    c.model = model
    c.n_neurons = 1
    c.n_exc = 1
    c.dt_ms = 0.5
    c.v_substeps = 1
    c.ring_slots = 2
    c.frac_bits = 32
    c.seed = 42
    c.lif_v_rest, c.lif_v_th, c.lif_v_reset = -60.0, -50.0, -65.0
    c.lif_tau_m, c.lif_t_ref_ms, c.lif_r = 8.0, 1.5, 2.0
    c.tau_exc_ms, c.tau_inh_ms, c.tau_nmda_ms = 2.0, 10.0, 100.0
    c.e_exc, c.e_inh = 0.0, -70.0
    c.gmax_ampa, c.gmax_nmda, c.gmax_gaba, c.mg_mM = 1.0, 0.5, 1.0, 1.0
    c.izh_a[0], c.izh_b[0], c.izh_c[0], c.izh_d[0] = 0.02, 0.2, -65.0, 8.0
    c.izh_a[1], c.izh_b[1], c.izh_c[1], c.izh_d[1] = 0.02, 0.25, -65.0, 2.0
    c.hh_g_na, c.hh_g_k, c.hh_g_l = 120.0, 36.0, 0.3
    c.hh_e_na, c.hh_e_k, c.hh_e_l, c.hh_c_m = 50.0, -77.0, -54.387, 1.0
    c.hh_spike_v = 0.0
    c.hh_spike_mode = abi.HH_SPIKE_REFERENCE_RESET
    c.hh_substeps = 1
    c.a_plus, c.a_minus, c.tau_plus_ms, c.tau_minus_ms = 0.01, 0.012, 20.0, 20.0
    c.w_min, c.w_max = 0.0, 10.0
    c.tau_y_ms, c.a3_plus = 100.0, 0.0003
    return c


def lif_network(n=10_000, seed=1, drive_scale=1.0, steps=10_000, plasticity=False,
                record_raster=False, raster_capacity=0, k_in=100):
    """Plain LIF: 80/20 E/I, fixed in-degree 100, delays 1-5 ms, Poisson 50 x 10 Hz x 0.3 mV."""
    c = base_config(abi.MODEL_LIF)
    c.n_neurons = n
    c.n_exc = (n * 4) // 5
    c.dt_ms = 0.1
    c.lif_v_rest, c.lif_v_reset, c.lif_v_th = -65.0, -65.0, -50.0
    c.lif_tau_m, c.lif_t_ref_ms, c.lif_r = 20.0, 2.0, 1.0
    c.tau_exc_ms, c.tau_inh_ms = 2.0, 10.0
    c.ring_slots = 51
    c.seed = seed
    c.drive = abi.DRIVE_POISSON
    c.drive_inputs = 50
    c.drive_rate_hz = 10.0
    c.drive_weight = 0.3 * drive_scale
    c.plasticity = abi.PLAST_PAIR if plasticity else abi.PLAST_NONE
    c.record_raster = 1 if record_raster else 0
    c.raster_capacity = raster_capacity
    return NetSpec(f"lif_{n}", c, k_in, (0.0, 0.5), (-2.0, 0.0), (10, 50), steps,
                   "LIF 80/20, in-degree 100, delays 1-5 ms, dt 0.1 ms")


def izh_stdp_network(n=1_000_000, seed=7, steps=10_000, plasticity=True, record_raster=False,
                     raster_capacity=0, k_in=100, rule="pair"):
    """Izhikevich + pair STDP: 80% RS / 20% FS, in-degree 100, delays 1-20 ms, N(0,5) per ms.
    rule="triplet" swaps in triplet STDP with TripletStdpRule's defaults (plasticity.rs:59-71)."""
    c = base_config(abi.MODEL_IZHIKEVICH)
    c.n_neurons = n
    c.n_exc = (n * 4) // 5
    c.dt_ms = 1.0
    c.v_substeps = 2
    c.izh_a[0], c.izh_b[0], c.izh_c[0], c.izh_d[0] = 0.02, 0.2, -65.0, 8.0
    c.izh_a[1], c.izh_b[1], c.izh_c[1], c.izh_d[1] = 0.1, 0.2, -65.0, 2.0
    c.izh_c_jitter = 15.0
    c.izh_d_jitter_frac = 0.75
    c.ring_slots = 21
    c.seed = seed
    c.drive = abi.DRIVE_GAUSSIAN
    c.drive_mean = 0.0
    c.drive_sigma = 5.0
    c.plasticity = abi.PLAST_PAIR if plasticity else abi.PLAST_NONE
    c.a_plus, c.a_minus, c.tau_plus_ms, c.tau_minus_ms = 0.01, 0.012, 20.0, 20.0
    c.w_min, c.w_max = 0.0, 10.0
    if plasticity and rule == "triplet":
        c.plasticity = abi.PLAST_TRIPLET
        c.tau_plus_ms, c.tau_minus_ms, c.tau_y_ms = 16.8, 33.7, 100.0
        c.a_plus, c.a_minus, c.a3_plus = 5e-4, 2e-4, 3e-4
    c.record_raster = 1 if record_raster else 0
    c.raster_capacity = raster_capacity
    return NetSpec(f"izh_stdp_{n}", c, k_in, (0.0, 6.0), (-5.0, 0.0), (1, 20), steps,
                   "Izhikevich RS/FS + pair STDP E->E, in-degree 100, delays 1-20 ms, dt 1 ms "
                   "(2 x 0.5 ms V sub-steps)")


def hh_network(n=200_000, seed=11, steps=100_000, record_raster=False, raster_capacity=0,
               k_in=100, w_exc=(0.0, 0.004), w_inh=(-0.02, 0.0), dt_ms=0.01, hh_substeps=1):
    """HH conductance network: AMPA/GABA-A/NMDA, OU drive (6, 3, 5 ms), dt 0.01 ms.

    ``hh_substeps`` > 1 integrates the membrane with that many Euler sub-steps of
    dt/hh_substeps per step, synaptic input and drive held over the step."""
    c = base_config(abi.MODEL_HH)
    c.n_neurons = n
    c.n_exc = (n * 4) // 5
    c.dt_ms = dt_ms
    c.hh_substeps = hh_substeps
    c.hh_spike_mode = abi.HH_SPIKE_CROSSING
    c.hh_spike_v = 0.0
    c.tau_exc_ms, c.tau_inh_ms, c.tau_nmda_ms = 2.0, 10.0, 100.0
    c.mg_mM = 1.0
    d_lo = int(round(1.0 / dt_ms))
    d_hi = int(round(2.0 / dt_ms))
    c.ring_slots = d_hi + 1
    c.seed = seed
    c.drive = abi.DRIVE_OU
    c.drive_mean, c.drive_sigma, c.drive_tau_ms = 6.0, 3.0, 5.0
    c.record_raster = 1 if record_raster else 0
    c.raster_capacity = raster_capacity
    return NetSpec(f"hh_{n}", c, k_in, w_exc, w_inh, (d_lo, d_hi), steps,
                   "HH + AMPA/GABA-A/NMDA(Mg 1 mM), OU 6/3/5 ms, delays 1-2 ms, dt 0.01 ms")


def hh_surrogate_network(n=200_000, seed=11, steps=20_000, record_raster=False,
                         raster_capacity=0, k_in=100, w_exc=(0.0, 0.004), w_inh=(-0.02, 0.0)):
    """The HH network stepped by the learned surrogate: dt 0.05 ms, one MLP evaluation per step
    standing for five held-input 0.01 ms Euler sub-steps (the plain step of
    hh_network(dt_ms=0.05, hh_substeps=5), which is also its per-neuron fallback)."""
    spec = hh_network(n=n, seed=seed, steps=steps, record_raster=record_raster,
                      raster_capacity=raster_capacity, k_in=k_in, w_exc=w_exc, w_inh=w_inh,
                      dt_ms=0.05, hh_substeps=5)
    spec.cfg.model = abi.MODEL_HH_SURROGATE
    spec.name = f"hh_surrogate_{n}"
    spec.description = ("HH + AMPA/GABA-A/NMDA, OU 6/3/5 ms, dt 0.05 ms, learned surrogate "
                        "5->64->64->4 for five 0.01 ms Euler sub-steps")
    return spec


# Training recipe of the surrogate: 50M transitions of the plain integrator (half under the
# network's OU drive, half under random drives), targets scaled by 1 / (|dV| + 0.5 mV), Adam with
# a cosine schedule over minibatches of 512 x 2 x 64 = 65,536 transitions.
SURROGATE_RECIPE = {"n_samples": 50_000_000, "data_seed": 3, "random_fraction": 0.5,
                    "scale_eps": 0.5, "steps": 20_000, "waves": 512, "chunks_per_wave": 2,
                    "lr": 2e-3, "lr_min": 1e-5, "seed": 5}


def _mix64(x):
    """SplitMix64 finaliser (Steele, Lea & Flood 2014): per-network seeds from a master seed."""
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


def population_base(record_raster=False, raster_capacity=0):
    """Shared neuron/drive settings of the batched Izhikevich population: RS/FS 80/20 with the
    c/d jitter of the 1M workload, dt 1 ms with two 0.5 ms V sub-steps, delays 1-10 ms (11 ring
    slots), and a 5 Hz Poisson drive of 100 inputs x 4 (current units) per network (the
    workload leaves the drive's size open)."""
    c = izh_stdp_network(n=2, plasticity=False).cfg
    c.ring_slots = 11
    c.drive = abi.DRIVE_POISSON
    c.drive_inputs = 100
    c.drive_rate_hz = 5.0
    c.drive_weight = 4.0
    c.drive_mean = 0.0
    c.drive_sigma = 0.0
    c.record_raster = 1 if record_raster else 0
    c.raster_capacity = raster_capacity
    return c


def population_nets(n_nets=256, seed=2026, n_range=(1000, 64000), k_range=(20, 200),
                    w_exc=(0.0, 3.0), w_inh=(-5.0, 0.0), delays=(1, 10)):
    """Networks of the batched population: N log-uniform in n_range, in-degree uniform in
    k_range, 80/20 E/I, own seed each (SplitMix64 of the master seed and the index)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), n_nets)).astype(np.int64)
    ks = rng.integers(k_range[0], k_range[1] + 1, n_nets)
    nets = []
    for i in range(n_nets):
        b = abi.NsxBatchNet()
        b.n_neurons = int(ns[i])
        b.n_exc = int(ns[i]) * 4 // 5
        b.k_in = int(ks[i])
        b.d_lo, b.d_hi = delays
        b.seed = _mix64((seed << 20) + i)
        b.w_exc_lo, b.w_exc_hi = w_exc
        b.w_inh_lo, b.w_inh_hi = w_inh
        nets.append(b)
    return nets


def population_single(base, net):
    """Config of one network of the population run on its own (the sequential reference)."""
    c = base.copy()
    c.n_neurons = net.n_neurons
    c.n_exc = net.n_exc
    c.seed = net.seed
    return c


def lif_large_network(n=100_000_000, seed=5, steps=10_000, plasticity=True, drive_scale=1.0):
    """Single-GPU large scale: the plain LIF network of lif_network (80/20, in-degree 100,
    delays 1-5 ms, Poisson 50 x 10 Hz x 0.3 mV, dt 0.1 ms) at 10^8 neurons / 10^10 synapses,
    with pair STDP on E->E (StdpRule defaults) and per-neuron traces."""
    spec = lif_network(n=n, seed=seed, drive_scale=drive_scale, steps=steps,
                       plasticity=plasticity)
    c = spec.cfg
    c.a_plus, c.a_minus, c.tau_plus_ms, c.tau_minus_ms = 0.01, 0.012, 20.0, 20.0
    c.w_min, c.w_max = 0.0, 10.0
    spec.name = f"lif_stdp_{n}"
    spec.description = "LIF 80/20, in-degree 100, delays 1-5 ms, dt 0.1 ms, pair STDP E->E"
    return spec
