"""The benchmark workloads: wiring rule, drive and step counts.

Neuron density is the one of the 300-neuron unit cube for every workload;
the connection radius is set for the wanted mean out-degree K in the bulk,
r = (K / (density * 4/3 * pi))^(1/3); neurons near the cube faces have fewer
neighbours, so the realised degree is reported, not assumed.
"""
import math

from ._lib import DriveSpec
from .engine import Network, default_params

DENSITY = 300.0


def radius_for_degree(k, density=DENSITY):
    return (k / (density * 4.0 / 3.0 * math.pi)) ** (1.0 / 3.0)


def side_for(n, density=DENSITY):
    return (n / density) ** (1.0 / 3.0)


# name -> wiring, drive, steps
WORKLOADS = {
    # 300 neurons in the unit cube, radius 0.15, delays 1..20, w ~ U[0.2, 0.8),
    # 6 inputs: two groups of 2 sharing a sine-modulated Bernoulli train
    # p = 0.05 (1 + sin(2 pi t / 200)) and 2 independent Bernoulli(0.05).
    "plain-300": dict(
        wiring=dict(n_neurons=300, side=1.0, radius=0.15, delay_min=1, delay_max=20, w_min=0.2, w_max=0.8),
        drive=dict(n_inputs=6, n_groups=2, n_correlated=4, period=200, p0=0.05, amp_corr=1.0, amp_ind=0.0),
        steps=10_000),
    # 20,000 neurons, ~50 synapses each, delays 1..32, 5% inputs in 4
    # correlated groups + 1 uncorrelated group, same Bernoulli/sine drive.
    "plain-20k": dict(
        wiring=dict(n_neurons=20_000, side=side_for(20_000), radius=radius_for_degree(50), delay_min=1,
                    delay_max=32, w_min=0.2, w_max=0.8),
        drive=dict(n_inputs=1000, n_groups=4, n_correlated=800, period=200, p0=0.05, amp_corr=1.0,
                   amp_ind=1.0),
        steps=50_000),
    # 10^6 neurons x ~100 synapses (10^8), delays 1..64, 1% inputs in 16
    # correlation groups at the same Bernoulli/sine drive (p0 = 0.05), network
    # mean rate p ~ 0.005 per step spread over the whole network.  Under the
    # reference's STDP a 100-synapse network is bistable: weights that let
    # threshold crossings carry ~0.005 ignite into p ~ 0.13 within a few
    # thousand steps (potentiation also weakens the inhibitory synapses).
    # Weights in [0.05, 0.3) with the reference's 20% inhibitory synapses
    # (src/NeuCor.cpp:473) stay below ignition, and the reference's background
    # fires (src/NeuCor.cpp:604-607, 1/600 per step at its defaults) at
    # 0.0045 per neuron-step give the 5 Hz: p = 0.0050 at the start, 0.0066
    # after 10^4 steps as STDP strengthens the weights.
    "plain-1m": dict(
        wiring=dict(n_neurons=1_000_000, side=side_for(1_000_000), radius=radius_for_degree(100), delay_min=1,
                    delay_max=64, w_min=0.05, w_max=0.3, inhib_frac=0.2),
        drive=dict(n_inputs=10_000, n_groups=16, n_correlated=10_000, period=200, p0=0.05, amp_corr=1.0,
                   amp_ind=0.0, bg_prob=0.0045),
        steps=10_000),
}


def batched_specs(count=4096, seed=0, n_range=(200, 20_000), k_range=(10, 60), steps=20_000):
    """The batched form's networks: neuron counts log-uniform in n_range,
    K (synapses per neuron in the bulk) uniform in k_range, each wired by the
    300-neuron workload's rule (density 300 per unit volume, radius for K,
    delays 1..20 steps proportional to distance, w ~ U[0.2, 0.8)); per-network
    seed and input-correlation amplitude uniform in [0, 1].  Inputs follow the
    300-neuron workload's drive scaled with N: 2% of the neurons (at least 6),
    two thirds of them in 2 sine-modulated groups at p = 0.05 (1 + amp
    sin(2 pi t / 200)), the rest independent Bernoulli(0.05).  Returns a list
    of dicts (wiring, drive, steps) for build_spec()."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        k = float(rng.uniform(*k_range))
        net_seed = int(rng.integers(0, 2**63 - 1))
        amp = float(rng.uniform(0.0, 1.0))
        n_in = max(6, int(round(0.02 * n)))
        n_corr = 2 * ((2 * n_in // 3) // 2)
        out.append(dict(
            wiring=dict(n_neurons=n, side=side_for(n), radius=radius_for_degree(k), delay_min=1, delay_max=20,
                        w_min=0.2, w_max=0.8, seed=net_seed),
            drive=dict(n_inputs=n_in, n_groups=2, n_correlated=n_corr, period=200, p0=0.05, amp_corr=amp,
                       amp_ind=0.0, seed=net_seed),
            steps=steps, k=k))
    return out


def build_spec(spec, n_threads=1):
    """(Network, DriveSpec, Params, steps) of one batched_specs() entry."""
    w = dict(spec["wiring"])
    seed = w.pop("seed")
    d = dict(spec["drive"])
    dseed = d.pop("seed")
    net = Network.generate(seed=seed, spatial_order=1, n_threads=n_threads, **w)
    return net, drive_spec(d, seed=dseed), default_params(), spec["steps"]


def drive_spec(d, seed=0, bg_prob=None):
    bg = d.get("bg_prob", 0.0) if bg_prob is None else bg_prob
    return DriveSpec(d["n_inputs"], d["n_groups"], d["n_correlated"], d["period"], d["p0"], d["amp_corr"],
                     d["amp_ind"], bg, seed)


def build(name, seed=0, n_threads=0, **override):
    """(Network, DriveSpec, Params, steps) of a named workload."""
    w = dict(WORKLOADS[name]["wiring"])
    w.update(override.pop("wiring", {}))
    d = dict(WORKLOADS[name]["drive"])
    d.update(override.pop("drive", {}))
    net = Network.generate(seed=seed, spatial_order=1, n_threads=n_threads, **w)
    return net, drive_spec(d, seed=seed), default_params(), override.get("steps", WORKLOADS[name]["steps"])
