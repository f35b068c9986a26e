/*
 * hs_models.h -- model identifiers, state-variable slots and parameter layouts
 * shared by the C-ABI (hs.h) and every caller.  Plain C, no device code.
 *
 * Each model is one Brian2 NeuronGroup equation set that the reference's code
 * generator would turn into a state-update kernel (brian2cuda/templates/
 * common_group.cu:108-160 via stateupdate.cu:1-3), a thresholder
 * (threshold.cu:19-45) and a resetter (reset.cu:3-27).  Parameters are SI
 * values as Brian2 stores them (volt, second, siemens, farad).
 */
#ifndef HS_MODELS_H
#define HS_MODELS_H

/* precision of state, weights and arithmetic */
#define HS_F64 64
#define HS_F32 32

/* step-loop forms */
#define HS_ENGINE_PLAIN 0 /* per-phase kernels, effects applied to the state in place */
#define HS_ENGINE_FUSED 1 /* one launch per step, effects accumulated and applied next step */

/* weight modes of a pathway */
#define HS_W_SCALAR 0  /* one weight for all synapses (on_pre='v += w' with constant w) */
#define HS_W_SYNAPSE 1 /* one weight per synapse */
#define HS_W_PRE 2     /* one weight per presynaptic neuron (batched parameter sweeps) */

/*
 * HS_MODEL_LIF_CUBA -- examples/cuba.py:122-133, method='exact':
 *   dv/dt  = (ge+gi-(v-El))/taum : volt (unless refractory)
 *   dge/dt = -ge/taue ; dgi/dt = -gi/taui
 *   threshold v>Vt, reset v=Vr, refractory
 * params: El, Vt, Vr, taum, taue, taui, refractory[, free]
 *   free (optional, default 0): 1 drops "(unless refractory)" -- v integrates
 *   while refractory and refractoriness only blocks the threshold, as in the
 *   STDP networks of brian2cuda/tests/test_cpp_cuda_consistency.py:41-44
 */
#define HS_MODEL_LIF_CUBA 1
#define HS_LIF_CUBA_NVARS 3
#define HS_LIF_CUBA_NPARAMS 7
#define HS_LIF_CUBA_NPARAMS_MAX 8
#define HS_V 0
#define HS_GE 1
#define HS_GI 2

/*
 * HS_MODEL_LIF_NOISE -- examples/brunelhakim.py:155-160, method='euler':
 *   dV/dt = (-V + mu + sigma*sqrt(tau)*xi)/tau : volt
 *   threshold V>theta, reset V=Vr, refractory
 * params: tau, mu, sigma, theta, Vr, refractory, noise_stream
 */
#define HS_MODEL_LIF_NOISE 2
#define HS_LIF_NOISE_NVARS 1
#define HS_LIF_NOISE_NPARAMS 7

/*
 * HS_MODEL_HH_COBA -- examples/cobahh.py:160-185 / brian2cuda/tests/features/
 * speed.py:73-113, method='exponential_euler':
 *   Hodgkin-Huxley (Traub-Miles rates) with conductance synapses ge, gi
 *   threshold v>v_th, refractory (no reset)
 * vars: v, m, n, h, ge, gi
 * params: Cm, gl, El, EK, ENa, gNa, gK, VT, Ee, Ei, taue, taui, v_th, refractory
 */
#define HS_MODEL_HH_COBA 3
#define HS_HH_NVARS 6
#define HS_HH_NPARAMS 14
#define HS_HH_V 0
#define HS_HH_M 1
#define HS_HH_N 2
#define HS_HH_H 3
#define HS_HH_GE 4
#define HS_HH_GI 5

/*
 * HS_MODEL_SPIKEGEN -- SpikeGeneratorGroup (brian2cuda/templates/spikegenerator.cu:43-80):
 *   neuron i spikes at step s for every (i, s) given to hs_net_set_spikegen;
 *   with a period of P steps the pattern repeats (step t fires the spikes of
 *   t mod P).  No state: one placeholder variable; no reset, no refractoriness.
 * params: period (seconds; 0: no repetition)
 */
#define HS_MODEL_SPIKEGEN 4
#define HS_SPIKEGEN_NVARS 1
#define HS_SPIKEGEN_NPARAMS 1

/*
 * HS_MODEL_POISSON -- PoissonGroup(N, rates) (Brian2's PoissonGroup: threshold
 * 'rand() < rates*dt', used as the input of examples/stdp.py:162 and
 * tests/features/speed.py:519,649): neuron i spikes at step t when
 * u(seed, stream, i, t) < rates_i * dt, u the 53-bit Philox uniform in [0, 1)
 * (hs_rand) and rates_i * dt computed in the group's precision.  No reset, no
 * refractoriness.
 * vars: rates (Hz; every neuron starts at the rate parameter)
 * params: rate (Hz) [, stream (Philox stream, default 0)]
 */
#define HS_MODEL_POISSON 5
#define HS_POISSON_NVARS 1
#define HS_POISSON_NPARAMS 1
#define HS_POISSON_NPARAMS_MAX 2
#define HS_POISSON_RATES 0

/*
 * HS_MODEL_CODE -- a group whose model is given as code (hs_net_add_group_code):
 * the statements of its state update, its threshold condition and its reset,
 * as the reference's code generator emits them for any NeuronGroup
 * (common_group.cu:108-160 via stateupdate.cu, threshold.cu:19-45,
 * reset.cu:3-27), compiled for the device when the network is built.  A
 * network's code models get the ids HS_MODEL_CODE + 0, 1, ... in the order
 * they are added, at most HS_MAX_CODE_MODELS of them.
 */
#define HS_MODEL_CODE 64
#define HS_MAX_CODE_MODELS 4
#define HS_IS_CODE_MODEL(m) ((m) >= HS_MODEL_CODE && (m) < HS_MODEL_CODE + HS_MAX_CODE_MODELS)

#define HS_MAX_VARS 8

/*
 * Spike-timing-dependent plasticity on a pathway (hs_net_pathway_stdp), the
 * synapse model of brian2cuda/tests/test_cpp_cuda_consistency.py:45-61:
 *   dApre/dt = -Apre/taupre (event-driven); dApost/dt = -Apost/taupost (event-driven); w
 *   on_pre : target += w*scale; Apre += dApre; w = clip(w + Apost, wmin, wmax)
 *   on_post: Apost += dApost;   w = clip(w + Apre, wmin, wmax)
 * Before either, both traces decay from the synapse's last update to now.
 * params: taupre, taupost, dApre, dApost, wmin, wmax, scale
 * synapse variables (hs_net_get_synapse_var): w, Apre, Apost, lastupdate (step)
 */
#define HS_STDP_NPARAMS 7
#define HS_SYN_W 0
#define HS_SYN_APRE 1
#define HS_SYN_APOST 2
#define HS_SYN_LASTUPDATE 3
/* synapse variables of a pathway with synaptic code (hs_net_pathway_code):
   variable k is HS_SYN_CODE + k, at most HS_MAX_SYN_VARS of them */
#define HS_SYN_CODE 16
#define HS_MAX_SYN_VARS 8

/* "never spiked" marker for the integer last-spike step */
#define HS_NEVER_SPIKED (-1073741824)

#endif /* HS_MODELS_H */
