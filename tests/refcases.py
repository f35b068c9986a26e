"""Small networks that restate the reference's spike-propagation tests.

The reference drives them with a SpikeGeneratorGroup and a NeuronGroup 'v:1'
(no dynamics), threshold v>1, reset v=0, on_pre 'v += 1.1'
(brian2cuda/tests/test_synaptic_propagations.py).  Here the input is a driver
neuron that starts above threshold (it spikes at step 0, once, or at every step
when its rest and reset sit above threshold), and 'v:1' is a CUBA LIF with an
infinite membrane time constant (v is constant between effects, exactly: Em
rounds to 1).  A target that receives 1.1 in the synapses slot of step d crosses
threshold at the update of step d+1, so it spikes at step delay+1, as the
reference's expectations state.
"""
import numpy as np

from brian2cuda_amd.networks import GroupSpec, NetSpec, PathSpec
from brian2cuda_amd.runtime import HS_MODEL_LIF_CUBA, HS_MODEL_SPIKEGEN, HS_W_SCALAR, HS_W_SYNAPSE

INF_TAU = 1e30  # exp(-dt/tau) rounds to 1.0: v is exactly constant


def static_params(refractory=0.0):
    """'v:1' with threshold v>1, reset v=0 (El=0, Vt=1, Vr=0)."""
    return [0.0, 1.0, 0.0, INF_TAU, 5e-3, 10e-3, refractory]


def driver(n=1, every_step=False):
    """Neurons that spike at step 0 only (every_step=False) or at every step."""
    if every_step:  # rest and reset at 2 > threshold, no refractoriness
        return GroupSpec(HS_MODEL_LIF_CUBA, n, [2.0, 1.0, 2.0, INF_TAU, 5e-3, 10e-3, 0.0], {0: 2.0, 1: 0.0, 2: 0.0})
    # refractory 1 s: after its spike at step 0 it stays silent for any of these runs
    return GroupSpec(HS_MODEL_LIF_CUBA, n, [0.0, 1.0, 0.0, INF_TAU, 5e-3, 10e-3, 1.0], {0: 2.0, 1: 0.0, 2: 0.0})


def push_outer_loop(dt=1e-4):
    """test_CudaSpikeQueue_push_outer_loop (test_synaptic_propagations.py:11-48):
    15*1025 targets, delays alternating 0 / dt; even targets spike at step 1, odd at step 2."""
    N = 15 * 1025
    g_in = driver(1)
    g = GroupSpec(HS_MODEL_LIF_CUBA, N, static_params(), {0: 0.0, 1: 0.0, 2: 0.0})
    pre = np.zeros(N, np.int32)
    post = np.arange(N, dtype=np.int32)
    d = (np.arange(N) % 2).astype(np.int32)
    p = PathSpec(0, 1, 0, pre, post, delay_steps=d, wmode=HS_W_SCALAR, wscalar=1.1)
    spec = NetSpec("push_outer_loop", dt, 1, [g_in, g], [p], spike_monitors=[1])
    expect = {1: np.arange(0, N, 2), 2: np.arange(1, N, 2)}
    return spec, 3, expect


def push_outer_loop2(dt=1e-4):
    """test_CudaSpikeQueue_push_outer_loop2 (test_synaptic_propagations.py:51-150):
    2048 targets; input 0 fires at every step with weight 0 and delays j; input 1
    fires once with weight 1.1 and delays j except delay[1024] = 1023.  Target j
    spikes at step delay_j + 1: 1023 and 1024 together, nothing at 1025."""
    N = 2048
    g_every = driver(1, every_step=True)
    g_once = driver(1)
    g = GroupSpec(HS_MODEL_LIF_CUBA, N, static_params(), {0: 0.0, 1: 0.0, 2: 0.0})
    post = np.arange(N, dtype=np.int32)
    d0 = np.arange(N, dtype=np.int32)
    d1 = d0.copy()
    d1[1024] = 1023
    paths = [
        PathSpec(0, 2, 0, np.zeros(N, np.int32), post, delay_steps=d0, wmode=HS_W_SYNAPSE, weights=np.zeros(N)),
        PathSpec(1, 2, 0, np.zeros(N, np.int32), post, delay_steps=d1, wmode=HS_W_SYNAPSE,
                 weights=np.full(N, 1.1)),
    ]
    spec = NetSpec("push_outer_loop2", dt, 1, [g_every, g_once, g], paths, spike_monitors=[2])
    expect = {}
    for j in range(N):
        expect.setdefault(int(d1[j]) + 1, []).append(j)
    expect = {k: np.array(v) for k, v in expect.items()}
    return spec, N + 2, expect


def circular_eventspaces(dt=1e-4):
    """test_circular_eventspaces (test_synaptic_propagations.py:153-181): homogeneous
    delays 0, 2dt, 5dt and heterogeneous j*dt for j=3,4; spike steps 1, 3, 6, 4, 5."""
    g_in = driver(1)
    g = GroupSpec(HS_MODEL_LIF_CUBA, 5, static_params(), {0: 0.0, 1: 0.0, 2: 0.0})
    z = np.zeros(1, np.int32)
    paths = [
        PathSpec(0, 1, 0, z, np.array([0], np.int32), scalar_delay=0, wscalar=1.1),
        PathSpec(0, 1, 0, z, np.array([1], np.int32), scalar_delay=2, wscalar=1.1),
        PathSpec(0, 1, 0, z, np.array([2], np.int32), scalar_delay=5, wscalar=1.1),
        PathSpec(0, 1, 0, np.zeros(2, np.int32), np.array([3, 4], np.int32),
                 delay_steps=np.array([3, 4], np.int32), wscalar=1.1),
    ]
    spec = NetSpec("circular_eventspaces", dt, 1, [g_in, g], paths, spike_monitors=[1])
    expect = {1: np.array([0]), 3: np.array([1]), 4: np.array([3]), 5: np.array([4]), 6: np.array([2])}
    return spec, 7, expect


def circular_eventspaces_spikegenerator(sem=0, dt=1e-4):
    """test_circular_eventspaces_spikegenerator (test_synaptic_propagations.py:185-229):
    a SpikeGeneratorGroup neuron fires at steps 0..11 into five targets with
    delays 0, 2, 5 (homogeneous) and 3, 4 (heterogeneous); each arrival adds 1.1.
    The reference runs it with resets scheduled before synapses, so a target
    spikes at every step d+1 .. d+12.  Brian2's default schedule (synapses, then
    resets) makes it spike every second step, d+1, d+3, .., d+11: the arrival
    of the spiking step is erased by the reset (brian2 issue 1332, cited there).
    Both semantics here restate the default schedule (ACC adds a step's effects
    at the start of the next step, before the deferred reset), so both expect
    every second step; sem is kept for symmetry with the other cases."""
    n = 12
    g_in = GroupSpec(HS_MODEL_SPIKEGEN, 1, [0.0], spikes=(np.zeros(n, np.int32), np.arange(n, dtype=np.int64)))
    g = GroupSpec(HS_MODEL_LIF_CUBA, 5, static_params(), {0: 0.0, 1: 0.0, 2: 0.0})
    z = np.zeros(1, np.int32)
    paths = [
        PathSpec(0, 1, 0, z, np.array([0], np.int32), scalar_delay=0, wscalar=1.1),
        PathSpec(0, 1, 0, z, np.array([1], np.int32), scalar_delay=2, wscalar=1.1),
        PathSpec(0, 1, 0, z, np.array([2], np.int32), scalar_delay=5, wscalar=1.1),
        PathSpec(0, 1, 0, np.zeros(2, np.int32), np.array([3, 4], np.int32),
                 delay_steps=np.array([3, 4], np.int32), wscalar=1.1),
    ]
    spec = NetSpec("circular_eventspaces_spikegenerator", dt, 1, [g_in, g], paths, spike_monitors=[1])
    expect = {}
    for j, d in enumerate([0, 2, 5, 3, 4]):
        steps = range(d + 1, d + n, 2)
        for t in steps:
            expect.setdefault(t, []).append(j)
    expect = {k: np.sort(np.array(v)) for k, v in expect.items()}
    return spec, n + 6, expect


def multiple_runs(delays, dt=1e-4):
    """test_multiple_runs_with_scalar_delay / _with_heterogeneous_delay
    (test_network_multiple_runs.py:12-31, :91-108): an input neuron spiking at
    every step (threshold 'True': a generator with a one-step period) adds 1 to
    'v : 1' of len(delays) targets through synapses of the given delays; a
    StateMonitor (when='start') records v over two runs of two steps.
    Returns the spec and the expected trace [target][step]."""
    g_in = GroupSpec(HS_MODEL_SPIKEGEN, 1, [dt], spikes=(np.zeros(1, np.int32), np.zeros(1, np.int64)))
    n = len(delays)
    params = static_params()
    params[1] = 1e30  # 'v : 1' without threshold
    g = GroupSpec(HS_MODEL_LIF_CUBA, n, params, {0: 0.0, 1: 0.0, 2: 0.0})
    p = PathSpec(0, 1, 0, np.zeros(n, np.int32), np.arange(n, dtype=np.int32),
                 delay_steps=np.array(delays, np.int32), wscalar=1.0)
    spec = NetSpec("multiple_runs", dt, 1, [g_in, g], [p], state_monitors=[(1, 0, 0, n)])
    expect = np.array([[max(0, t - d) for t in range(4)] for d in delays], dtype=np.float64)
    return spec, expect


def effects_cancel(n_pre=15360, n_post=64, dt=1e-4):
    """test_synaptic_effect_modes (test_synaptic_propagations.py:316-340) restated
    with two on_pre pathways: every pre neuron spikes at every step, one pathway
    adds 1 and the other subtracts 1 from the same targets, so v stays 0."""
    g_in = driver(n_pre, every_step=True)
    g = GroupSpec(HS_MODEL_LIF_CUBA, n_post, static_params(), {0: 0.0, 1: 0.0, 2: 0.0})
    rng = np.random.default_rng(3)
    post = rng.integers(0, n_post, n_pre).astype(np.int32)
    pre = np.arange(n_pre, dtype=np.int32)
    paths = [PathSpec(0, 1, 0, pre, post, scalar_delay=1, wscalar=1.0),
             PathSpec(0, 1, 0, pre, post, scalar_delay=1, wscalar=-1.0)]
    spec = NetSpec("effects_cancel", dt, 1, [g_in, g], paths, spike_monitors=[1])
    return spec, 20, {}


def spikes_by_step(i, t):
    out = {}
    for step in np.unique(t):
        out[int(step)] = np.sort(i[t == step])
    return out


def bundle_prefs(dt=1e-4):
    """test_threads_per_synapse_bundle_prefs (test_synaptic_propagations.py:343-374):
    a SpikeGeneratorGroup neuron fires once at step 0 into targets 3 and 4 of
    five static neurons with delays j*dt; each arrival adds 1.1, so neuron 3
    spikes at step 4 and neuron 4 at step 5 over a 6-step run, whatever the
    launch shapes (the reference varies the threads per bundle)."""
    g_in = GroupSpec(HS_MODEL_SPIKEGEN, 1, [0.0], spikes=(np.zeros(1, np.int32), np.zeros(1, np.int64)))
    g = GroupSpec(HS_MODEL_LIF_CUBA, 5, static_params(), {0: 0.0, 1: 0.0, 2: 0.0})
    p = PathSpec(0, 1, 0, np.zeros(2, np.int32), np.array([3, 4], np.int32),
                 delay_steps=np.array([3, 4], np.int32), wscalar=1.1)
    spec = NetSpec("bundle_prefs", dt, 1, [g_in, g], [p], spike_monitors=[1])
    return spec, 6, {4: np.array([3]), 5: np.array([4])}


def ring(N=1000, dt=1e-4, w=0.5e-3):
    """test_cuda_standalone (test_cuda_standalone.py:16-49): dV/dt = (-40 mV - V)/tau
    (unless refractory), tau 1 ms, threshold V > -50 mV, reset V = -60 mV,
    refractory 5 ms, V(0) = -i mV; on_pre V += w (0.5 mV) onto the neighbours
    |i - j| < 5, i != j, delay 0; 100 ms.  The reference expects 17,000 to 18,000
    spikes, the first at t = 0."""
    mV, ms = 1e-3, 1e-3
    g = GroupSpec(HS_MODEL_LIF_CUBA, N, [-40 * mV, -50 * mV, -60 * mV, 1 * ms, 5 * ms, 10 * ms, 5 * ms],
                  {0: -np.arange(N) * mV, 1: 0.0, 2: 0.0})
    i = np.repeat(np.arange(N), 9)
    j = np.repeat(np.arange(N), 9) + np.tile(np.arange(-4, 5), N)
    ok = (j >= 0) & (j < N) & (j != i)
    p = PathSpec(0, 0, 0, i[ok].astype(np.int32), j[ok].astype(np.int32), scalar_delay=0, wscalar=w)
    spec = NetSpec("ring", dt, 1, [g], [p], spike_monitors=[0])
    return spec, 1000


def ring_numpy(N=1000, steps=1000, dt=1e-4, w=0.5e-3):
    """The ring restated independently in numpy on Brian2's schedule (state
    update with refractoriness timestep(t - lastspike) >= timestep(5 ms), threshold,
    synapses with delay 0, reset), FP64.  Returns (spike count, first spike step)."""
    mV = 1e-3
    E = np.exp(-dt / 1e-3)
    V = -np.arange(N) * mV
    last = np.full(N, -10 ** 9)
    n, first = 0, None
    for s in range(steps):
        nr = (s - last) >= 50
        V = np.where(nr, -40 * mV + (V + 40 * mV) * E, V)
        spk = np.nonzero((V > -50 * mV) & nr)[0]
        if first is None and spk.size:
            first = s
        n += spk.size
        last[spk] = s
        add = np.zeros(N)
        for d in (-4, -3, -2, -1, 1, 2, 3, 4):
            j = spk + d
            np.add.at(add, j[(j >= 0) & (j < N)], w)
        V = V + add
        V[spk] = -60 * mV
    return n, first


def example_run(N=100, dt=1e-4):
    """brian2cuda's example_run (brian2cuda/__init__.py:40-77): dv/dt = -v/tau,
    tau 10 ms, threshold v > 1, reset v = 0, refractory 5 ms, v(0) = i/100, run
    1 ms; the reference asserts v == arange(N)/N * exp(-1 ms/tau)."""
    ms = 1e-3
    # taue, taui: unused (ge = gi = 0), but they must differ from taum
    g = GroupSpec(HS_MODEL_LIF_CUBA, N, [0.0, 1.0, 0.0, 10 * ms, 5 * ms, 20 * ms, 5 * ms],
                  {0: np.arange(N) / 100.0, 1: 0.0, 2: 0.0})
    spec = NetSpec("example_run", dt, 1, [g], [], spike_monitors=[0])
    return spec, 10, np.arange(N) / N * np.exp(-1 * ms / (10 * ms))
