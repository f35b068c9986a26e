"""CPU pins of the oracle's multicompartment solver (the reference's
section-wise Thomas + coupling + combine algorithm) against solves it does not
share: a dense solve of the compartment/branch-point system, a banded solve of
an unbranched cable, and Rall's closed-form steady state.  (No GPU.)"""
import numpy as np
from scipy import linalg

import oracle
from brian2cuda_amd import morphology as mo


def _dense_step(tb, t, v, m, n, h, I):
    """One implicit-Euler step of tree t, unknowns = compartments + branch points."""
    c0, c1 = tb.tree_comp[t], tb.tree_comp[t + 1]
    s0, s1 = tb.tree_sec[t], tb.tree_sec[t + 1]
    N, ns = c1 - c0, s1 - s0
    area, rl1, rl2, Cm = tb.area[c0:c1], tb.rl1[c0:c1], tb.rl2[c0:c1], tb.Cm[c0:c1]
    gl, El, gNa, ENa, gK, EK = tb.chan
    gna, gk = gNa * m[c0:c1] ** 3 * h[c0:c1], gK * n[c0:c1] ** 4
    gtot = gl + gna + gk
    I0 = gl * El + gna * ENa + gk * EK + I[c0:c1] / area
    M = np.zeros((N + ns + 1, N + ns + 1))
    rhs = np.zeros(N + ns + 1)

    def couple(a, b, g, area_a=None, area_b=None):
        # current g (x_b - x_a) into a (divided by its area when a is a compartment)
        for x, y, ar in ((a, b, area_a), (b, a, area_b)):
            s = 1.0 / ar if ar is not None else 1.0
            M[x, x] += g * s
            M[x, y] -= g * s
    for j in range(N):
        M[j, j] += Cm[j] / tb.dt + gtot[j]
        rhs[j] = Cm[j] / tb.dt * v[c0 + j] + I0[j]
    for s in range(ns):
        f = tb.sec_first[s0 + s]
        l = f + tb.sec_n[s0 + s] - 1
        p = tb.sec_parent[s0 + s]
        start_node = N + (0 if p < 0 else p + 1)
        end_node = N + s + 1
        couple(f, start_node, rl1[f] / tb.Ri, area[f], None)
        couple(l, end_node, rl2[l] / tb.Ri, area[l], None)
        for j in range(f + 1, l + 1):
            g = 1.0 / (tb.Ri * (1 / rl2[j - 1] + 1 / rl1[j]))
            couple(j, j - 1, g, area[j], area[j - 1])
    return np.linalg.solve(M, rhs)[:N]


def test_hines_step_matches_dense_solve():
    tb = mo.random_trees(n_trees=6, seed=11, cmin=20, cmax=60)
    ref = oracle.RefMC(tb, 64)
    rng = np.random.default_rng(1)
    v0 = 0.02 * rng.standard_normal(tb.n_comp)
    ref.set_state(0, v0)
    m, n, h, I = tb.init[1], tb.init[2], tb.init[3], tb.init[4]
    ref.run(1)
    v1 = ref.get_state(0)
    for t in range(tb.n_trees):
        want = _dense_step(tb, t, v0, m, n, h, I)
        got = v1[tb.tree_comp[t]:tb.tree_comp[t + 1]]
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-13)


def test_unbranched_cable_matches_banded_solve():
    tb = mo.cable(n_comp=50, I=0.05e-9)
    ref = oracle.RefMC(tb, 64)
    ref.run(1)
    v1 = ref.get_state(0)
    N = tb.n_comp
    gl = tb.chan[0]
    g = 1.0 / (tb.Ri * (1 / tb.rl2[:-1] + 1 / tb.rl1[1:]))  # between j and j+1 (sealed ends: no end terms)
    a = tb.area
    diag = tb.Cm / tb.dt + gl
    diag[:-1] += g / a[:-1]
    diag[1:] += g / a[1:]
    ab = np.zeros((3, N))
    ab[0, 1:] = -g / a[:-1]
    ab[1] = diag
    ab[2, :-1] = -g / a[1:]
    rhs = tb.Cm / tb.dt * tb.init[0] + gl * tb.chan[1] + tb.init[4] / a
    want = linalg.solve_banded((1, 1), ab, rhs)
    np.testing.assert_allclose(v1, want, rtol=1e-11, atol=1e-15)


def test_rall_cable_steady_state():
    """Passive sealed cable, current into one end: v(x) - El = r_a lam I cosh((L-x)/lam)/sinh(L/lam)
    (examples/compartmental/rall_cuda.py:63-76); compartments of lam/100."""
    L, d, I, n = 1000e-6, 1e-6, 0.02e-9, 200
    tb = mo.cable(n_comp=n, length=L, diameter=d, I=I, gl=1.0, El=0.0)
    ref = oracle.RefMC(tb, 64)
    ref.run(8000)  # 200 ms = 20 membrane time constants
    v = ref.get_state(0)
    rm = 1 / (tb.chan[0] * np.pi * d)
    ra = 4 * tb.Ri / (np.pi * d ** 2)
    lam = np.sqrt(rm / ra)
    x = (np.arange(n) + 0.5) * (L / n)
    want = ra * lam * I * np.cosh((L - x) / lam) / np.sinh(L / lam)
    np.testing.assert_allclose(v, want, rtol=0.02)


def test_random_trees_are_depth_first():
    tb = mo.random_trees(n_trees=20, seed=2)
    for t in range(tb.n_trees):
        s0, s1 = tb.tree_sec[t], tb.tree_sec[t + 1]
        par = tb.sec_parent[s0:s1]
        assert par[0] == -1 and np.all(par[1:] < np.arange(1, s1 - s0))
        assert np.all(np.bincount(par[1:] , minlength=s1 - s0) <= 2)
        n = tb.tree_comp[t + 1] - tb.tree_comp[t]
        assert 100 <= n <= 420
        assert tb.sec_first[s0] == 0 and tb.sec_first[s1 - 1] + tb.sec_n[s1 - 1] == n


def test_hh_tree_spikes_and_fp32_channels_track_fp64():
    tb = mo.random_trees(n_trees=8, seed=4)
    a, b = oracle.RefMC(tb, 64), oracle.RefMC(tb, 32)
    a.run(2000)
    b.run(2000)
    ca, cb = a.spike_counts(), b.spike_counts()
    assert ca.sum() > 0
    assert np.abs(ca - cb).max() <= 1
    assert np.all(np.isfinite(a.get_state(0)))
