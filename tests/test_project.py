"""CPU checks of the standalone project writer (brian2cuda_amd.project): the
written project holds the spec's arrays, compiles against libhs.so, and its
binary handles the reference's command line (templates/main.cu:29-62,
objects.cu:89-183) -- unknown variables, unreadable files and a missing device
end it with exit code 1 and a message, before any device work."""
import json
import os
import subprocess

import numpy as np
import pytest

from brian2cuda_amd import networks as nw
from brian2cuda_amd import project

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "brian2cuda_amd", "libhs.so")
needs_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="libhs.so not built (run __graft_entry__.build())")


def test_project_holds_the_spec_arrays(tmp_path):
    spec = nw.cuba(N=300, seed=3, record=20)
    info = project.write_project(spec, tmp_path, 100, precision=32, engine="fused", options={0: 128})
    sa = tmp_path / "static_arrays"
    np.testing.assert_array_equal(np.fromfile(sa / "_array_neurongroup_v", "<f8"), spec.groups[0].init[0])
    for k, name in enumerate(["synapses", "synapses_1"]):
        np.testing.assert_array_equal(np.fromfile(sa / f"_dynamic_array_{name}__synaptic_pre", "<i4"),
                                      spec.paths[k].pre)
        np.testing.assert_array_equal(np.fromfile(sa / f"_dynamic_array_{name}__synaptic_post", "<i4"),
                                      spec.paths[k].post)
    on_disk = json.loads((tmp_path / "project.json").read_text())
    assert on_disk == json.loads(json.dumps(info))
    assert info["precision"] == 32 and info["engine"] == 1 and info["steps"] == 100
    assert [m["kind"] for m in info["monitors"]] == ["spike", "state"]
    src = (tmp_path / "main.cpp").read_text()
    assert "hs_net_set_option(net, 0, 128)" in src and "hs_net_run(net, 100LL)" in src
    assert "hs_net_create(32, 1," in src


def test_names_are_checked(tmp_path):
    spec = nw.cuba(N=100, seed=1)
    with pytest.raises(ValueError, match="unique"):
        project.write_project(spec, tmp_path, 10, names={"groups": ["synapses"]})
    with pytest.raises(ValueError, match="identifier"):
        project.write_project(spec, tmp_path, 10, names={"groups": ["my group"]})
    with pytest.raises(ValueError, match="names for"):
        project.write_project(spec, tmp_path, 10, names={"paths": ["only_one"]})


@needs_lib
def test_binary_command_line_errors(tmp_path):
    spec = nw.cuba(N=200, seed=2)
    project.write_project(spec, tmp_path, 10)
    exe = project.build_project(tmp_path)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")

    def run(*args):
        return subprocess.run([exe, *args], cwd=tmp_path, capture_output=True, text=True, env=env, timeout=60)

    r = run("neurongroup.bogus=1")
    assert r.returncode == 1 and "Cannot set unknown variable 'neurongroup.bogus'" in r.stderr
    r = run("neurongroup.v=no_such_file.dat")
    assert r.returncode == 1 and "Could not read 'no_such_file.dat'" in r.stderr
    np.zeros(199).tofile(tmp_path / "short.dat")  # one value short
    r = run("neurongroup.v=short.dat")
    assert r.returncode == 1 and "expected 1600 bytes" in r.stderr
    r = run("neurongroup.v=1.5x")
    assert r.returncode == 1 and "Cannot parse value" in r.stderr
    r = run("noequals")
    assert r.returncode == 1 and "Cannot parse argument" in r.stderr
    # valid overrides get as far as the device, which is absent here: a clean error
    np.full(200, -0.055).tofile(tmp_path / "v.dat")
    r = run("--results_dir", str(tmp_path / "res"), "neurongroup.v=v.dat", "neurongroup.ge=TRUE")
    assert r.returncode == 1 and "hs_init(0) failed" in r.stderr
    assert not (tmp_path / "res").exists()


@needs_lib
def test_every_model_and_pathway_kind_compiles(tmp_path):
    """The writer's output for each built-in model, weight mode, delay form,
    device connect, STDP, spike generator and monitor kind is valid C++."""
    specs = [nw.cuba(N=100, seed=1, record=5), nw.cobahh(N=100, seed=1), nw.brunel(N=200, K=20, seed=1),
             nw.batch_cuba(n_nets=3, seed=1, nmin=50, nmax=80), nw.stdp_poisson_example(N=200, K=20),
             nw.brunel(N=200, K=20, seed=1, device_connect=True, fixed_outdegree=True),
             nw.code_synapses(npo=50, nl=10), nw.code_mix(N=300, seed=1), nw.stdp_feedforward(2, 20, 10, code=True),
             nw.cuba_conditional(N=200, seed=1)]
    specs[0].rate_monitors = [(0, 10, 60)]
    for k, spec in enumerate(specs):
        d = tmp_path / f"p{k}"
        project.write_project(spec, d, 5)
        assert os.path.exists(project.build_project(d))
    assert "hs_net_connect_fixed(net, g_neurongroup, g_neurongroup, 0, 0, 200, 0, 200, 20," in \
        (tmp_path / f"p{len(specs) - 5}" / "main.cpp").read_text()
    assert "hs_net_pathway_code(net, p_synapses, 2, sn," in (tmp_path / f"p{len(specs) - 4}" / "main.cpp").read_text()
    main = (tmp_path / f"p{len(specs) - 3}" / "main.cpp").read_text()
    assert main.count("hs_net_add_group_code(") == 3 and '"gr"' in main
    assert "hs_net_pathway_code_post(net, p_synapses," in (tmp_path / f"p{len(specs) - 2}" / "main.cpp").read_text()
    main = (tmp_path / f"p{len(specs) - 1}" / "main.cpp").read_text()
    assert main.count("hs_net_connect_code(") == 2
    assert '"((bool)(((i != j))) && (bool)(((hs_cg_pymod(abs((i - j)), 41LL) == 0LL))))", "", 0.7, 0,' in main
