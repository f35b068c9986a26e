"""GPU: the standalone project path end to end -- a network written as a
project (main.cpp + static arrays), compiled against libhs.so, run as a
separate binary with the reference's command line, and its result files read
back -- gives the oracle's records bit for bit."""
import numpy as np
import pytest

from brian2cuda_amd import networks as nw
from brian2cuda_amd import project

pytestmark = pytest.mark.gpu


def _oracle(oracle_lib, spec, prec, sem, steps):
    ref, rsm, rst = oracle_lib.build_ref(spec, prec, sem)
    ref.run(steps)
    return ref, rsm, rst


@pytest.mark.parametrize("prec,engine", [(64, "plain"), (32, "fused")])
def test_c1_project_matches_oracle(hs, oracle_lib, tmp_path, prec, engine):
    """C1 (CUBA, 4000 neurons, spikes of all, v of 0..99 every step) for 1000 steps."""
    spec = nw.cuba(N=4000, seed=11)
    steps = 1000
    print(f"project C1 prec={prec} engine={engine}", flush=True)
    project.write_project(spec, tmp_path, steps, precision=prec, engine=engine)
    project.build_project(tmp_path)
    project.run_project(tmp_path, "results", timeout=120)
    res = project.read_results(tmp_path)
    ref, rsm, rst = _oracle(oracle_lib, spec, prec, oracle_lib.SEQ if engine == "plain" else oracle_lib.ACC, steps)
    ri, rstep = ref.spikes(rsm[0])
    s = res["spikemonitor"]
    assert s["N"] == len(ri) > 0
    np.testing.assert_array_equal(s["i"], ri)
    np.testing.assert_array_equal(s["t"], rstep.astype(np.float64) * spec.dt)
    np.testing.assert_array_equal(s["count"], np.bincount(ri, minlength=4000))
    v = res["statemonitor"]
    assert v["N"] == steps and v["v"].shape == (steps, 100)
    np.testing.assert_array_equal(v["t"], np.arange(steps) * spec.dt)
    np.testing.assert_array_equal(v["indices"], np.arange(100))
    np.testing.assert_array_equal(v["v"].astype(np.float64), ref.state_monitor(rst[0], 100))
    for var, name in enumerate(["v", "ge", "gi"]):
        np.testing.assert_array_equal(res[f"neurongroup.{name}"], ref.get_state(0, var))
    wall, frac = res["last_run_info"]
    assert wall > 0 and frac == 1.0


def test_command_line_overrides_initial_state(hs, oracle_lib, tmp_path):
    """owner.var=file and owner.var=value set the initial state the run starts
    from (objects.cu:89-183); the oracle started from the same values agrees."""
    spec = nw.cuba(N=2000, seed=5, record=0)
    project.write_project(spec, tmp_path, 500)
    project.build_project(tmp_path)
    rng = np.random.default_rng(9)
    v0 = -0.060 + 0.012 * rng.random(2000)
    v0.tofile(tmp_path / "v0.dat")
    project.run_project(tmp_path, tmp_path / "out", args=["neurongroup.v=v0.dat", "neurongroup.ge=0.001"])
    res = project.read_results(tmp_path, tmp_path / "out")
    spec.groups[0].init = {0: v0, 1: 0.001, 2: 0.0}
    ref, rsm, _ = _oracle(oracle_lib, spec, 64, oracle_lib.SEQ, 500)
    ri, rstep = ref.spikes(rsm[0])
    assert len(ri) > 0
    np.testing.assert_array_equal(res["spikemonitor"]["i"], ri)
    np.testing.assert_array_equal(res["spikemonitor"]["t"], rstep * spec.dt)
    np.testing.assert_array_equal(res["neurongroup.v"], ref.get_state(0, 0))


def test_device_connect_and_noise_project(hs, oracle_lib, tmp_path):
    """Brunel-Hakim (connect(p) generated on the device, drawn delays, Philox
    noise) through the project: the oracle regenerates the same synapses."""
    spec = nw.brunel(N=5000, K=200, seed=21, device_connect=True, monitors=True)
    project.write_project(spec, tmp_path, 600, precision=32, engine="fused")
    project.build_project(tmp_path)
    project.run_project(tmp_path, "results", timeout=120)
    res = project.read_results(tmp_path)
    ref, rsm, _ = _oracle(oracle_lib, spec, 32, oracle_lib.ACC, 600)
    ri, rstep = ref.spikes(rsm[0])
    assert len(ri) > 0
    np.testing.assert_array_equal(res["spikemonitor"]["i"], ri)
    np.testing.assert_array_equal(res["spikemonitor"]["t"], rstep * spec.dt)
    np.testing.assert_array_equal(res["neurongroup.V"], ref.get_state(0, 0).astype(np.float32).astype(np.float64))


def test_fixed_outdegree_project(hs, oracle_lib, tmp_path):
    """The fixed out-degree connect through the project (hs_net_connect_fixed in
    main.cpp): spikes and state equal the oracle's."""
    spec = nw.brunel(N=3000, K=300, seed=23, device_connect=True, fixed_outdegree=True, monitors=True)
    project.write_project(spec, tmp_path, 500, precision=32, engine="fused")
    project.build_project(tmp_path)
    project.run_project(tmp_path, "results", timeout=120)
    res = project.read_results(tmp_path)
    ref, rsm, _ = _oracle(oracle_lib, spec, 32, oracle_lib.ACC, 500)
    ri, rstep = ref.spikes(rsm[0])
    assert len(ri) > 0
    np.testing.assert_array_equal(res["spikemonitor"]["i"], ri)
    np.testing.assert_array_equal(res["spikemonitor"]["t"], rstep * spec.dt)
    np.testing.assert_array_equal(res["neurongroup.V"], ref.get_state(0, 0).astype(np.float32).astype(np.float64))


def test_conditional_connect_project(hs, oracle_lib, tmp_path):
    """connect with a probability expression and a condition through the
    project (hs_net_connect_code in main.cpp: the binary compiles them for the
    device when it runs): spikes and state equal the oracle's."""
    spec = nw.cuba_conditional(N=4000, seed=13)
    project.write_project(spec, tmp_path, 500, precision=64, engine="plain")
    project.build_project(tmp_path)
    project.run_project(tmp_path, "results", timeout=120)
    res = project.read_results(tmp_path)
    ref, rsm, _ = _oracle(oracle_lib, spec, 64, oracle_lib.SEQ, 500)
    ri, rstep = ref.spikes(rsm[0])
    assert len(ri) > 100
    np.testing.assert_array_equal(res["spikemonitor"]["i"], ri)
    np.testing.assert_array_equal(res["spikemonitor"]["t"], rstep * spec.dt)
    np.testing.assert_array_equal(res["neurongroup.v"], ref.get_state(0, 0))


def test_described_cuba_example_through_the_device_layer(hs, oracle_lib, tmp_path):
    """examples/cuba.py described object by object (equation text, threshold,
    reset, namespace, connect('i<Ne', p=0.02), on_pre 'ge += we') goes through
    the device's mapping, the project writer and the binary; the oracle run of
    the same spec (connectivity regenerated by its restatement) agrees."""
    from brian2cuda_amd import device as dv
    mV, ms = 1e-3, 1e-3
    ns = {"taum": 20 * ms, "taue": 5 * ms, "taui": 10 * ms, "Vt": -50 * mV, "Vr": -60 * mV, "El": -49 * mV,
          "we": 1.62 * mV, "wi": -9 * mV, "Ne": 3200}
    eqs = ("dv/dt  = (ge+gi-(v-El))/taum : volt (unless refractory)\n"
           "dge/dt = -ge/taue : volt\ndgi/dt = -gi/taui : volt\n")
    v0 = -60 * mV + np.random.default_rng(3).random(4000) * 10 * mV
    g = dv.GroupDesc("neurongroup", 4000, equations=eqs, threshold="v>Vt", reset="v = Vr", refractory=5 * ms,
                     method="exact", namespace=ns, init={"v": v0, "ge": 0.0, "gi": 0.0})
    syns = [dv.SynapsesDesc("synapses", "neurongroup", "neurongroup", "ge += we", namespace=ns, p=0.02,
                            condition="i<Ne", delay=0.1 * ms, seed=5),
            dv.SynapsesDesc("synapses_1", "neurongroup", "neurongroup", "gi += wi", namespace=ns, p=0.02,
                            condition="i>=Ne", delay=0.1 * ms, seed=5)]
    mons = [dv.MonitorDesc("spikemonitor", "spike", "neurongroup"),
            dv.MonitorDesc("statemonitor", "state", "neurongroup", variable="v", record=range(50))]
    spec, names = dv.network_spec([g], syns, mons, dt=0.1 * ms, seed=5, name="cuba_example")
    project.write_project(spec, tmp_path, 800, names=names)
    project.build_project(tmp_path)
    project.run_project(tmp_path, "results", timeout=120)
    res = project.read_results(tmp_path)
    ref, rsm, rst = _oracle(oracle_lib, spec, 64, oracle_lib.SEQ, 800)
    ri, rstep = ref.spikes(rsm[0])
    assert len(ri) > 100
    np.testing.assert_array_equal(res["spikemonitor"]["i"], ri)
    np.testing.assert_array_equal(res["spikemonitor"]["t"], rstep * spec.dt)
    np.testing.assert_array_equal(res["statemonitor"]["v"], ref.state_monitor(rst[0], 50))


def test_code_models_through_the_project(hs, oracle_lib, tmp_path):
    """Three code models and a Poisson input (networks.code_mix) through the
    project: main.cpp passes the generated statements to hs_net_add_group_code;
    spikes and final state equal the oracle's."""
    spec = nw.code_mix(N=2000, seed=9)
    project.write_project(spec, tmp_path, 800, precision=64, engine="fused")
    project.build_project(tmp_path)
    project.run_project(tmp_path, "results", timeout=180)
    res = project.read_results(tmp_path)
    ref, rsm, _ = _oracle(oracle_lib, spec, 64, oracle_lib.ACC, 800)
    names = ["spikemonitor", "spikemonitor_1", "spikemonitor_2"]
    for k, name in enumerate(names):
        ri, rstep = ref.spikes(rsm[k])
        assert len(ri) > 0
        np.testing.assert_array_equal(res[name]["i"], ri)
        np.testing.assert_array_equal(res[name]["t"], rstep * spec.dt)
    np.testing.assert_array_equal(res["neurongroup.v"], ref.get_state(0, 0))
    np.testing.assert_array_equal(res["neurongroup.w"], ref.get_state(0, 1))


def test_synaptic_code_through_the_project(hs, oracle_lib, tmp_path):
    """Synaptic code with per-synapse initial values (networks.code_synapses)
    through the project: spikes and the target state equal the oracle's."""
    spec = nw.code_synapses()
    project.write_project(spec, tmp_path, 1000, precision=64, engine="plain")
    project.build_project(tmp_path)
    project.run_project(tmp_path, "results", timeout=180)
    res = project.read_results(tmp_path)
    ref, rsm, _ = _oracle(oracle_lib, spec, 64, oracle_lib.SEQ, 1000)
    ri, rstep = ref.spikes(rsm[0])
    assert len(ri) > 0
    np.testing.assert_array_equal(res["spikemonitor"]["i"], ri)
    np.testing.assert_array_equal(res["spikemonitor"]["t"], rstep * spec.dt)
    np.testing.assert_array_equal(res["neurongroup_1.v"], ref.get_state(1, 0))
    np.testing.assert_array_equal(res["neurongroup_1.ge"], ref.get_state(1, 1))


def test_stdp_as_code_through_the_project(hs, oracle_lib, tmp_path):
    """STDP written as synaptic code (event-driven traces, on_pre, on_post)
    through the project: spikes and voltages equal the oracle's."""
    spec = nw.stdp_feedforward(3, code=True)
    project.write_project(spec, tmp_path, 2000, precision=64, engine="fused")
    project.build_project(tmp_path)
    project.run_project(tmp_path, "results", timeout=180)
    res = project.read_results(tmp_path)
    ref, rsm, _ = _oracle(oracle_lib, spec, 64, oracle_lib.ACC, 2000)
    ri, rstep = ref.spikes(rsm[0])
    assert len(ri) > 0
    np.testing.assert_array_equal(res["spikemonitor"]["i"], ri)
    np.testing.assert_array_equal(res["spikemonitor"]["t"], rstep * spec.dt)
    np.testing.assert_array_equal(res["neurongroup_1.v"], ref.get_state(1, 0))
