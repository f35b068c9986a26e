"""PoissonGroup (HS_MODEL_POISSON) in the oracle: Brian2's threshold
'rand() < rates*dt' with the Philox uniform (pinned by the Random123
known-answer vectors).  Checked by the statistics the definition fixes: the
spike count is Binomial(N*steps, rates*dt), first-spike steps are Geometric,
rates of 0 never fire and rates*dt >= 1 fires every step; per-neuron rates and
the stream parameter act as stated."""
import numpy as np
import pytest
from scipy import stats

import oracle
from brian2cuda_amd.runtime import HS_MODEL_POISSON

DT = 1e-4


def _run(params, n, steps, prec=64, rates=None, seed=7):
    net = oracle.RefNet(prec, oracle.SEQ, DT, seed)
    g = net.add_group(HS_MODEL_POISSON, n, params)
    if rates is not None:
        net.set_state(g, 0, rates)
    m = net.add_spike_monitor(g)
    net.run(steps)
    i, t = net.spikes(m)
    net.close()
    return i, t


@pytest.mark.parametrize("prec", [64, 32])
def test_count_and_first_spike_distribution(prec):
    n, steps, rate = 20000, 1000, 20.0
    p = rate * DT
    i, t = _run([rate], n, steps, prec)
    mean, sd = n * steps * p, np.sqrt(n * steps * p * (1 - p))
    assert abs(len(i) - mean) < 5 * sd
    # first spike of each neuron that fired: Geometric(p) on 0, 1, 2, ... truncated at steps
    first = np.full(n, -1)
    o = np.lexsort((t, i))
    i, t = i[o], t[o]
    starts = np.r_[True, i[1:] != i[:-1]]
    first[i[starts]] = t[starts]
    f = first[first >= 0]
    cdf = lambda k: (1 - (1 - p) ** (np.floor(k) + 1)) / (1 - (1 - p) ** steps)
    ks = stats.kstest(f + 0.5, cdf)  # +0.5: compare at the middle of each integer step
    assert ks.pvalue > 1e-3
    # ISIs inside the run: a pair of consecutive spikes k steps apart occurs at
    # (steps - k) places with probability p^2 (1-p)^(k-1) each
    same = i[1:] == i[:-1]
    isi = np.diff(t)[same]
    k = np.arange(1, steps)
    wgt = (steps - k) * (1 - p) ** (k - 1)
    m1 = (k * wgt).sum() / wgt.sum()
    sd = np.sqrt((k * k * wgt).sum() / wgt.sum() - m1 ** 2)
    assert abs(isi.mean() - m1) < 5 * sd / np.sqrt(isi.size)
    assert abs(isi.size - n * (p * p * wgt.sum())) < 5 * np.sqrt(n * p * p * wgt.sum())


def test_zero_and_certain_rates():
    i, _ = _run([0.0], 500, 300)
    assert i.size == 0
    i, t = _run([1.0 / DT], 50, 20)  # rates*dt = 1 > every uniform in [0, 1)
    assert i.size == 50 * 20
    assert np.array_equal(np.bincount(t, minlength=20), np.full(20, 50))


def test_per_neuron_rates_and_stream():
    n, steps = 4000, 2000
    rates = np.where(np.arange(n) < n // 2, 0.0, 50.0)
    i, _ = _run([10.0], n, steps, rates=rates)
    assert not np.any(i < n // 2)
    c = np.bincount(i, minlength=n)[n // 2:]
    lam = steps * 50.0 * DT
    assert abs(c.mean() - lam) < 5 * np.sqrt(lam / c.size)
    a = _run([30.0, 0.0], 1000, 500)
    b = _run([30.0, 1.0], 1000, 500)
    assert not (np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]))
    assert np.array_equal(_run([30.0, 1.0], 1000, 500)[0], b[0])


def test_bad_parameters():
    net = oracle.RefNet(64, oracle.SEQ, DT, 1)
    with pytest.raises(Exception):
        net.add_group(HS_MODEL_POISSON, 10, [-1.0])
    with pytest.raises(Exception):
        net.add_group(HS_MODEL_POISSON, 10, [1.0, 0.0, 0.0])
    net.close()
