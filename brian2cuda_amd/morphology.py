"""Multicompartment workloads: batches of branched cylinder trees.

A tree is a list of sections (Brian2's morphology sections) numbered depth-first
so that a parent precedes its children; each section is a cylinder of n
compartments of equal length.  The flat arrays follow Brian2's SpatialNeuron
variables (area, r_length_1, r_length_2, Cm per compartment; section start /
parent tables) so that the prep of the reference (spatialstateupdate.cu:18-93)
applies unchanged:
  area        = pi d l                    (m^2, l = compartment length)
  r_length_1  = r_length_2 = pi d^2 / (2 l)   (m; cross-section over half length)

Channels: Hodgkin-Huxley (1952) Na/K with rest at 0 V
(examples/compartmental/hh_with_spikes_cuda.py:20-40), SI units per m^2.
"""
from dataclasses import dataclass, field

import numpy as np

um, nA, ms, mV = 1e-6, 1e-9, 1e-3, 1e-3

# gl, El, gNa, ENa, gK, EK  (S/m^2, V)
HH1952 = [0.3e-3 * 1e4, 10.613 * mV, 120e-3 * 1e4, 115 * mV, 36e-3 * 1e4, -12 * mV]
CM_DEFAULT = 1e-2   # 1 uF/cm^2 in F/m^2
RI_DEFAULT = 1.0    # 100 ohm cm = 1 ohm m
SPIKE_VTH = 50 * mV


@dataclass
class TreeBatch:
    dt: float
    tree_comp: np.ndarray      # int32[n_trees+1]
    tree_sec: np.ndarray       # int32[n_trees+1]
    sec_first: np.ndarray      # int32[n_sections]: first compartment (tree-local)
    sec_n: np.ndarray          # int32[n_sections]
    sec_parent: np.ndarray     # int32[n_sections]: tree-local parent section, -1 for the root
    area: np.ndarray           # float64[n_comp]
    rl1: np.ndarray
    rl2: np.ndarray
    Cm: np.ndarray
    Ri: float = RI_DEFAULT
    chan: list = field(default_factory=lambda: list(HH1952))
    vth: float = SPIKE_VTH
    init: dict = field(default_factory=dict)   # 0 v, 1 m, 2 n, 3 h, 4 I (point current, A)

    @property
    def n_trees(self):
        return len(self.tree_comp) - 1

    @property
    def n_comp(self):
        return int(self.tree_comp[-1])

    @property
    def n_sections(self):
        return int(self.tree_sec[-1])


def _random_tree(rng, target, nmin=5, nmax=15):
    """Sections (n_comp, parent) of a random binary tree with about `target` compartments."""
    n = [int(rng.integers(nmin, nmax + 1))]
    parent = [-1]
    children = {0: []}
    leaves = [0]
    while sum(n) < target:
        s = leaves.pop(int(rng.integers(len(leaves))))
        for _ in range(2):
            k = len(n)
            n.append(int(rng.integers(nmin, nmax + 1)))
            parent.append(s)
            children[k] = []
            children[s].append(k)
            leaves.append(k)
    # depth-first renumbering: a parent precedes its children, subtrees are contiguous
    order, stack = [], [0]
    while stack:
        s = stack.pop()
        order.append(s)
        stack.extend(reversed(children[s]))
    new = {old: i for i, old in enumerate(order)}
    n2 = [n[o] for o in order]
    p2 = [(-1 if parent[o] < 0 else new[parent[o]]) for o in order]
    return n2, p2


def random_trees(n_trees=2048, seed=1234, cmin=100, cmax=400, dmin=0.5 * um, dmax=2.0 * um, lmin=10 * um,
                 lmax=50 * um, imin=0.1 * nA, imax=1.0 * nA, dt=0.025 * ms):
    """C5: independent branched HH cells of cmin..cmax compartments (seeded)."""
    rng = np.random.default_rng(seed)
    tree_comp, tree_sec = [0], [0]
    sec_first, sec_n, sec_parent = [], [], []
    area, rl, Ipt = [], [], []
    for _ in range(n_trees):
        target = int(rng.integers(cmin, cmax + 1))
        n, parent = _random_tree(rng, target)
        first = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int32)
        sec_first.extend(first)
        sec_n.extend(n)
        sec_parent.extend(parent)
        I = np.zeros(int(sum(n)))
        I[0] = rng.uniform(imin, imax)
        for k in range(len(n)):
            d = rng.uniform(dmin, dmax)
            l = rng.uniform(lmin, lmax)
            area.extend([np.pi * d * l] * n[k])
            rl.extend([np.pi * d * d / (2 * l)] * n[k])
        Ipt.extend(I)
        tree_comp.append(tree_comp[-1] + int(sum(n)))
        tree_sec.append(tree_sec[-1] + len(n))
    nc = tree_comp[-1]
    tb = TreeBatch(dt, np.array(tree_comp, np.int32), np.array(tree_sec, np.int32), np.array(sec_first, np.int32),
                   np.array(sec_n, np.int32), np.array(sec_parent, np.int32), np.array(area), np.array(rl),
                   np.array(rl), np.full(nc, CM_DEFAULT))
    tb.init = {0: np.zeros(nc), 1: np.full(nc, 0.05), 2: np.full(nc, 0.32), 3: np.full(nc, 0.6), 4: np.array(Ipt)}
    return tb


def cable(n_comp=200, length=1000 * um, diameter=1 * um, I=0.02 * nA, gl=1.0, El=0.0, dt=0.025 * ms):
    """One unbranched passive cylinder (sealed ends, current into compartment 0)."""
    l = length / n_comp
    a = np.pi * diameter * l
    r = np.pi * diameter ** 2 / (2 * l)
    tb = TreeBatch(dt, np.array([0, n_comp], np.int32), np.array([0, 1], np.int32), np.array([0], np.int32),
                   np.array([n_comp], np.int32), np.array([-1], np.int32), np.full(n_comp, a), np.full(n_comp, r),
                   np.full(n_comp, r), np.full(n_comp, CM_DEFAULT), chan=[gl, El, 0.0, 0.0, 0.0, 0.0])
    Ipt = np.zeros(n_comp)
    Ipt[0] = I
    tb.init = {0: np.full(n_comp, El), 1: np.zeros(n_comp), 2: np.zeros(n_comp), 3: np.zeros(n_comp), 4: Ipt}
    return tb


def build_hs(tb, precision=64, options=None):
    """Create a runtime.Hines batch from a TreeBatch."""
    from .runtime import Hines
    h = Hines(tb, precision)
    for k, v in (options or {}).items():
        h.set_option(k, v)
    for var, val in tb.init.items():
        h.set_state(var, val)
    return h
