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
