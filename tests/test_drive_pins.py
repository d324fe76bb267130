"""The oracle's input drive and background fires, pinned by their
probabilities rather than by a second copy of their formula.

The drive stands in for InputFirer (src/NeuCor.cpp:305-345) with the
workloads' sine-modulated Bernoulli trains: group g fires at step s with
probability p(s) = p0 (1 + amp sin(2 pi (s mod T) / T)), all members of a
group together, different groups independently.  Background fires stand in
for the reference's per-step random scheduleFire (:604-607, :658-661): each
neuron is forced with probability bg_prob per step and still respects the
refractory window (:683-684).  The GPU's drive equals the oracle's draw for
draw (the parity tests compare rasters), so these checks pin both.

Networks here have no synapses, so every fire is a drive or background fire
and nothing else moves the potentials out of their rest/AP cycle."""
import math

import numpy as np
import pytest

T = 200  # the workloads' period in steps


def empty_net(n):
    return {"in_ptr": np.zeros(n + 1, np.int64), "in_src": np.zeros(0, np.uint32), "w": np.zeros(0, np.float32),
            "delay": np.zeros(0, np.uint8), "inh": np.zeros(0, np.uint8)}


def fire_matrix(O, n, steps, ids, groups, n_corr, p0, amp_corr, amp_ind, bg=0.0, seed=11, period=T):
    """Fires as a (steps, n) boolean matrix, row s-1 for step s."""
    o = O.Sim(empty_net(n))
    o.set_drive(ids, groups, n_corr, period, p0, amp_corr, amp_ind, bg, seed)
    s, i = o.step(steps, cap=n * steps + 1)
    m = np.zeros((steps, n), bool)
    m[s.astype(np.int64) - 1, i.astype(np.int64)] = True
    return m


def sine_fit(rate_by_phase, period=T):
    """Least-squares a + b sin(2 pi k / T) + c cos(2 pi k / T)."""
    k = np.arange(period)
    X = np.stack([np.ones(period), np.sin(2 * np.pi * k / period), np.cos(2 * np.pi * k / period)], 1)
    coef, *_ = np.linalg.lstsq(X, rate_by_phase, rcond=None)
    return coef


def test_per_phase_rate_follows_the_sine(oracle_mod):
    """64 correlated singleton groups at amp 0.8 and 64 independent ones at
    amp 0: over 500 periods the per-phase rate fits p0 + p0 amp sin(2 pi k/T)
    with no cosine term.  A flipped sign, a wrong period or a phase off by
    one step (cosine term p0 amp sin(2 pi / T) = 0.0025, about 15 sigma)
    fails."""
    n, p0, amp = 128, 0.1, 0.8
    steps = 500 * T
    ids = np.arange(n, dtype=np.uint32)
    m = fire_matrix(oracle_mod, n, steps, ids, ids, 64, p0, amp, 0.0)
    phase = np.arange(1, steps + 1) % T
    for cols, a_exp in ((slice(0, 64), amp), (slice(64, 128), 0.0)):
        per = np.bincount(phase, weights=m[:, cols].sum(1), minlength=T) / (64 * steps / T)
        a, b, c = sine_fit(per)
        sig = math.sqrt(p0 * (1 - p0) / (64 * steps))  # std of the mean rate
        assert abs(a - p0) < 5 * sig, (a, p0)
        assert abs(b - p0 * a_exp) < 5 * math.sqrt(2) * sig, (b, p0 * a_exp)
        assert abs(c) < 5 * math.sqrt(2) * sig, c
    # the largest single-phase deviation stays binomial
    per = np.bincount(phase, weights=m[:, :64].sum(1), minlength=T) / (64 * steps / T)
    k = np.arange(T)
    p_k = p0 * (1 + amp * np.sin(2 * np.pi * k / T))
    sd = np.sqrt(p_k * (1 - p_k) / (64 * steps / T))
    assert np.all(np.abs(per - p_k) < 5 * sd)


def test_threshold_special_cases(oracle_mod):
    """p = 0 never fires; p = 1/2 fires exactly when the Philox word is below
    2^31 (floor(p 2^32) with no rounding); p >= 1 always fires."""
    n, steps, seed = 8, 4000, 5
    ids = np.arange(n, dtype=np.uint32)
    assert not fire_matrix(oracle_mod, n, steps, ids, ids, n, 0.0, 0.0, 0.0, seed=seed).any()
    m = fire_matrix(oracle_mod, n, 400, ids, ids, n, 0.5, 0.0, 0.0, seed=seed)
    key = (seed & 0xFFFFFFFF, seed >> 32)
    want = np.array([[int(oracle_mod.philox((s, g, 4, 0), key)[0]) < (1 << 31) for g in range(n)]
                     for s in range(1, 401)])
    assert np.array_equal(m, want)
    assert fire_matrix(oracle_mod, n, steps, ids, ids, n, 1.0, 0.0, 0.0, seed=seed).all()
    # p0 = 0.6, amp = 1: p(k) >= 1 wherever sin >= 2/3, and those phases fire every time
    m = fire_matrix(oracle_mod, n, 20 * T, ids, ids, n, 0.6, 1.0, 0.0, seed=seed)
    s = np.arange(1, 20 * T + 1)
    sure = 0.6 * (1 + np.sin(2 * np.pi * (s % T) / T)) >= 1.0
    assert m[sure].all()
    never = 0.6 * (1 + np.sin(2 * np.pi * (s % T) / T)) <= 0.0
    assert not m[never].any()


def test_group_members_fire_together_and_groups_independently(oracle_mod):
    """4 groups of 4 members share their trains bit for bit; different groups'
    trains are uncorrelated (|r| below 5 / sqrt(steps))."""
    n, steps = 16, 40_000
    ids = np.arange(n, dtype=np.uint32)
    groups = (ids // 4).astype(np.uint32)
    m = fire_matrix(oracle_mod, n, steps, ids, groups, 4, 0.2, 0.0, 0.0)
    for g in range(4):
        blk = m[:, 4 * g:4 * g + 4]
        assert (blk == blk[:, :1]).all()
    tr = m[:, ::4].astype(np.float64)
    r = np.corrcoef(tr.T)
    off = r[~np.eye(4, dtype=bool)]
    assert np.all(np.abs(off) < 5 / math.sqrt(steps)), off
    # independent inputs (groups past n_corr with amp 0) behave the same way
    m2 = fire_matrix(oracle_mod, n, steps, ids, ids, 0, 0.2, 0.0, 0.0)
    r2 = np.corrcoef(m2.T.astype(np.float64))
    assert np.all(np.abs(r2[~np.eye(n, dtype=bool)]) < 5 / math.sqrt(steps))


@pytest.mark.parametrize("p", [1.0 / 600.0, 0.3, 1.0])
def test_background_rate_with_refractory_window(oracle_mod, p):
    """A neuron forced with probability p per step that may not fire within
    2 ms (2 steps at dt = 1 ms) of its last fire is a renewal process with
    mean interval 2 + 1/p steps: rate p / (1 + 2p).  At p = 1 it fires
    exactly every third step."""
    n, steps = 1000, 20_000 if p < 0.01 else 4000
    ids = np.zeros(0, np.uint32)
    m = fire_matrix(oracle_mod, n, steps, ids, ids, 0, 0.0, 0.0, 0.0, bg=p)
    want = p / (1 + 2 * p)
    # burn-in of the renewal process: start counting after 100 steps
    got = m[100:].mean()
    trials = n * (steps - 100)
    # a renewal count's variance is at most the binomial one here
    assert abs(got - want) < 5 * math.sqrt(want * (1 - want) / trials) + 1e-12, (got, want)
    if p == 1.0:
        assert np.array_equal(np.flatnonzero(m[:, 0]), np.arange(0, steps, 3))


def test_background_draws_are_per_neuron_and_step(oracle_mod):
    """Background fires of different neurons are independent (pairwise |r|
    small); after a fire the next two steps are refractory, and the first
    step past the window is a fresh draw: P(fire at s+3 | fire at s) = p."""
    n, steps, p = 64, 30_000, 0.05
    ids = np.zeros(0, np.uint32)
    m = fire_matrix(oracle_mod, n, steps, ids, ids, 0, 0.0, 0.0, 0.0, bg=p)
    r = np.corrcoef(m.T.astype(np.float64))
    assert np.all(np.abs(r[~np.eye(n, dtype=bool)]) < 5 / math.sqrt(steps))
    assert not (m[1:] & m[:-1]).any() and not (m[2:] & m[:-2]).any()
    nf = m[:-3].sum()
    cond = (m[3:] & m[:-3]).sum() / nf
    assert abs(cond - p) < 5 * math.sqrt(p * (1 - p) / nf), (cond, p)


def test_background_event_list_matches_the_simulation(oracle_mod):
    """The raw background draws that bench.py scripts into the reference's
    run() are the simulation's: on a network without synapses and a
    background probability of 1/2, the simulation fires at step s exactly the
    drawn neurons that are out of their refractory window."""
    n, steps, p, seed = 50, 300, 0.5, 9
    ids = np.zeros(0, np.uint32)
    m = fire_matrix(oracle_mod, n, steps, ids, ids, 0, 0.0, 0.0, 0.0, bg=p, seed=seed)
    es, ei = oracle_mod.bg_events(seed, p, n, 0, steps)
    drawn = np.zeros((steps, n), bool)
    drawn[es.astype(np.int64) - 1, ei.astype(np.int64)] = True
    assert abs(drawn.mean() - p) < 0.02
    last = np.full(n, -10)
    for s in range(steps):
        want = drawn[s] & (s - last > 2)
        assert np.array_equal(m[s], want), s
        last[want] = s
