/*
 * C-ABI of libmonki_hip.so: MONKI photon transport on one gfx950 GPU.
 *
 * Counterpart of the reference Metal bridge ABI v2 (kernels/metal/bridge.m:26-150), which
 * python/monki/execution/metal.py:125-156 binds with ctypes. Plain pointers and sizes only.
 * A handle owns the device buffers of one prepared scene. One caller per handle: calls on a
 * handle must not overlap. Every call is synchronous and returns after the device is done.
 *
 * Failure flags reported in error text (same meanings as kernels/metal/transport.metal:420-457):
 *   1 nonfinite probability, 2 detector traversal limit, 4 direction sampling limit,
 *   8 invalid Stokes vector, 16 history step limit; and 32 nonfinite tally accumulated on the
 *   device (monki_hip_batch_accumulate, no reference counterpart).
 */
#ifndef MONKI_HIP_H
#define MONKI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ABI version (2); a Python host refuses a library whose version differs.
 * Replaces monki_metal_abi_version (kernels/metal/bridge.m:26-28). */
int monki_hip_abi_version(void);

/* Validate packed float32 tables (layout v2, monki/execution/hip_tables.py), select the
 * device, upload tables, and reserve tally banks: tally_count floats per round (all shards)
 * for round_capacity rounds. Returns NULL and fills error on failure.
 * Replaces monki_metal_create (kernels/metal/bridge.m:30-84); no kernel source: the
 * kernels are compiled for gfx950 ahead of time. */
void *monki_hip_create(uint32_t device, const float *tables, size_t table_count,
                       size_t tally_count, uint32_t round_capacity, char *error,
                       size_t error_size);

/* Run one group of rounds. controls[8] = photon count of this dispatch, first flattened
 * photon, first round (1-based), seed, step limit, profile flag, photons per round,
 * rounds in the group. Tallies are zeroed, traced in float32, reduced over shards in
 * float64 on the device, and copied to out: rounds x (tally_count / shards) doubles, per round
 * [Stokes | paths | weights], or with the scattering decomposition (table header word 29)
 * [Stokes | Stokes_sing | Stokes_mult | contrib_sing | contrib_mult]. The tables select the
 * transport: nmat 1, 3 or 4, forward or backward (header word 28), detector at TOA or at the
 * ground (header word 30), result grid at the detector or on the other level (word 31).
 * Returns 0, or nonzero with error text (the round group is then rejected).
 * Replaces monki_metal_run (kernels/metal/bridge.m:88-133), whose out is float per shard. */
int monki_hip_run(void *handle, const uint32_t *controls, double *out, char *error,
                  size_t error_size);

/* monki_hip_create + monki_hip_run of one round group in one call that allocates nothing: the
 * packed tables (host, as for monki_hip_create) and controls in, the float64 tallies out in
 * monki_hip_run's order. No reference counterpart: for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid allocation).
 * Everything besides the caller's buffers lives in a 64 MiB device array of the library,
 * one per device for the life of the process; a round group needing more (tables + float32
 * banks of all shards and rounds + float64 reduced and output tallies) is rejected with an
 * error. Runs on the device's null stream and returns when done. Calls that overlap on one
 * device race on that array; calls on different devices do not. Returns 0, or nonzero with
 * error text. */
int monki_hip_run_static(uint32_t device, const float *tables, size_t table_count, const uint32_t *controls,
                         double *out, char *error, size_t error_size);

/* Same as monki_hip_run, with transport arithmetic in float64 on the same Philox streams.
 * No reference counterpart: the FP64 build the FP32 error is measured against. */
int monki_hip_run_fp64(void *handle, const uint32_t *controls, double *out, char *error,
                       size_t error_size);

/* Per-photon records of the first min(record_count, controls[0]) photons of a dispatch
 * (8 floats each: atmospheric events, surface events, end status, final weight, sum of
 * Stokes-I detector contributions times mu0/mu, sum of detector weights, sum of weighted
 * path contributions, last layer). precision is 32 or 64. Tallies are also returned as
 * monki_hip_run does. No reference counterpart: used to check histories photon by photon. */
int monki_hip_trace(void *handle, const uint32_t *controls, uint32_t precision, double *out,
                    float *records, size_t record_count, char *error, size_t error_size);

/* Event counters of the last run (monotonic since create): 0 photons, 1 atmospheric events,
 * 2 surface events, 3 free-path cell crossings, 4 explicit-equivalent detector cells,
 * 5 phase-table entries read, 6 rejection-sampling attempts, 7 tally atomics issued.
 * Copies min(count, 8) values. No reference counterpart (the reference counts nothing). */
int monki_hip_counters(void *handle, uint64_t *out, size_t count);

/* Accumulated device time of transport dispatches when the profile flag was set
 * (kernels/metal/bridge.m:121-123, 135-138). */
double monki_hip_gpu_seconds(void *handle);

/* Seconds of the transport kernel alone in the last run, from HIP events on its stream. */
double monki_hip_kernel_seconds(void *handle);

/* Layer paths each lane of a plain float32 launch keeps in LDS: the largest nz | 1 of the
 * session's scenes, or fewer when whole rows would leave less than three waves per SIMD (the
 * upper layers' paths then live in global memory). A diagnostic with no reference counterpart. */
int monki_hip_layer_rows(void *handle);

/* Device name of the handle (kernels/metal/bridge.m:140-143). */
const char *monki_hip_device_name(void *handle);

/* Device-helper probes for tests (the reference's probe_kernel technique,
 * tests/integration/test_metal.py:23-40, as fixed entry points). No reference counterpart.
 * monki_hip_probe_scatter: per case 8 doubles (uo, u1, phio, phi1, depol, I, Q, U) in,
 * 12 doubles out (cos_sca, C1, S1, C2, S2, F11/4, F22/4, F33/4, F12/4, I', Q', U') computed in
 * float32 as transport does for a Rayleigh event. monki_hip_probe_philox: per case 6 words
 * (counter[4], key[2]) in, 4 words out. Return 0 on success. */
int monki_hip_probe_scatter(uint32_t device, const double *cases, size_t count, double *out);
int monki_hip_probe_philox(uint32_t device, const uint32_t *cases, size_t count, uint32_t *out);

/* Release device buffers; NULL is accepted (kernels/metal/bridge.m:145-150). */
void monki_hip_close(void *handle);

/* ---- Batched scenes. No reference counterpart (the reference runs one scene per session).
 * A batch session holds many scenes (each its own packed tables, grid size, geometry and tally
 * shards), traced together by one persistent launch per round: waves start spread over the
 * scenes and move to the next scene when theirs has no photons left. */

/* tables: the packed float32 buffers of scene_count scenes back to back; scene i occupies
 * words offsets[i] .. offsets[i + 1]. block: lanes per block (64, 128 or 256). One round slot
 * per scene. NULL and error on failure; release with monki_hip_close. */
void *monki_hip_batch_create(uint32_t device, const float *tables, const size_t *offsets,
                             uint32_t scene_count, uint32_t block, char *error, size_t error_size);

/* One round (1-based index `round`) of every scene, photons[i] photons for scene i. out receives,
 * scene after scene, each scene's tallies in the monki_hip_run order (float64). use_graph: run
 * the round as a HIP graph captured on first use (clears, transport, reductions; the per-scene
 * controls are read from device memory), else launch the same sequence directly. */
int monki_hip_batch_run(void *handle, uint32_t round, uint32_t seed, uint32_t max_steps,
                        const uint32_t *photons, double *out, int use_graph, char *error,
                        size_t error_size);

/* As monki_hip_batch_run, but the layer-path and detector-weight tallies stay on the device:
 * per scene with layer paths, the round is added to float64 sums of the outputs and to the
 * online (Welford) moments of the AMF estimator, and only the Stokes tallies of each scene
 * (nmat x nx x ny x views, Fortran order, scene after scene) are copied to stokes.
 * photons[i] = 0 skips scene i this round (nothing traced or accumulated). */
int monki_hip_batch_accumulate(void *handle, uint32_t round, uint32_t seed, uint32_t max_steps,
                               const uint32_t *photons, double *stokes, int use_graph, char *error,
                               size_t error_size);

/* The device accumulators of one scene (with layer paths): count = 1 + P + 3 W + 3 L values,
 * P = nx ny nz views, W = nx ny views, L = nx ny (nz - 1) views:
 * [rounds | path_sum (P) | weight_sum (W) | path_mean, path_m2, covariance (L each) |
 *  weight_mean, weight_m2 (W each)], arrays in Fortran order. Zeros before the first
 * accumulating round. */
int monki_hip_batch_moments(void *handle, uint32_t scene, double *out, size_t count, char *error,
                            size_t error_size);

/* ---- Learned proposal (guided tracer). No reference counterpart: the reference samples
 * directions from the phase function only (mc_routines_general_mod.f90:492-932).
 * A proposal handle owns a BF16 MLP (12 -> 4 x 128 SiLU -> 40: an 8-component von
 * Mises-Fisher mixture over the outgoing direction), its float32 master parameters and Adam
 * state, and a device buffer of training samples. */

/* Number of float32 master parameters (56,360). */
size_t monki_hip_proposal_parameter_count(void);

/* Create a proposal on a device with seeded He initialization; NULL and error on failure. */
void *monki_hip_proposal_create(uint32_t device, uint32_t seed, char *error, size_t error_size);

/* Replace or read the float32 master parameters (count = parameter count). */
int monki_hip_proposal_set_parameters(void *proposal, const float *params, size_t count,
                                      char *error, size_t error_size);
int monki_hip_proposal_get_parameters(void *proposal, float *params, size_t count);

/* Reserve room for `events` training samples and for the per-photon totals of dispatches of up
 * to `photons` photons; drops the current samples. */
int monki_hip_proposal_reserve(void *proposal, size_t events, size_t photons, char *error,
                               size_t error_size);

/* Replace the samples: features [count][12], unit directions [count][3], weights [count]. */
int monki_hip_proposal_set_samples(void *proposal, const float *features, const float *directions,
                                   const float *weights, size_t count, char *error,
                                   size_t error_size);

/* Path target of the training weights (no reference counterpart): with mix > 0, logging also
 * records, per event, the layer paths the photon still travels before each later estimate,
 * weighted by the estimate's detector weight and by layer_weights[k] (one per layer of the
 * scene, at least nz), and a sample's training weight becomes (1 - mix) G + mix s P, G the
 * detector weight after the event, P that path weight, s = sum G / sum P over the dispatch.
 * mix 0 turns it off. Returns nonzero (message in error) for mix outside [0, 1] or no weights. */
int monki_hip_proposal_set_path_target(void *proposal, const float *layer_weights, size_t layers, float mix,
                                       char *error, size_t error_size);

/* Sample count and sum of sample weights; copy the first `count` samples back. */
int monki_hip_proposal_samples(void *proposal, uint64_t *count, double *weight_sum);
int monki_hip_proposal_read_samples(void *proposal, float *features, float *directions,
                                    float *weights, size_t count);

/* Gradient (float32, parameter order) and value of -sum w log q(direction) / sum w over the
 * first `count` samples. */
int monki_hip_proposal_gradients(void *proposal, size_t count, float *grad, double *loss,
                                 char *error, size_t error_size);

/* `steps` Adam steps over random minibatches of `batch` samples; losses[steps] receives each
 * step's minibatch loss (may be NULL). */
int monki_hip_proposal_train(void *proposal, uint32_t steps, uint32_t batch, float learning_rate,
                             uint64_t seed, double *losses, char *error, size_t error_size);

/* Network outputs [count][40] for features [count][12] (the path transport uses). */
int monki_hip_proposal_forward(void *proposal, const float *features, size_t count, float *out,
                               char *error, size_t error_size);

void monki_hip_proposal_close(void *proposal);

/* Attach a proposal to a transport session (NULL detaches) with the probability alpha of
 * drawing directions from the phase function, for the scattering events of the kinds in
 * species_mask (bit 0 gas, bit s + 1 tabulated species s) up to scattering order max_order;
 * other events are sampled as in monki_hip_run. The proposal must outlive the attachment. */
int monki_hip_attach_proposal(void *handle, void *proposal, float alpha, uint32_t species_mask,
                              uint32_t max_order, char *error, size_t error_size);

/* Guided tracer: the scattering orders (1 .. orders) at which the network's mixture is part of
 * the proposal; at higher orders the view lobes alone (the network does not run). The default
 * is every order. No reference counterpart (the reference has no learned proposal). */
int monki_hip_set_network_orders(void *handle, uint32_t orders, char *error, size_t error_size);

/* monki_hip_run with the attached proposal: directions from the phase function with
 * probability alpha, else from the mixture q; weights times p / (alpha p + (1 - alpha) q). */
int monki_hip_run_guided(void *handle, const uint32_t *controls, double *out, char *error,
                         size_t error_size);

/* Detector-directed sampling, no network: directions from the phase function with
 * probability alpha, else from von Mises-Fisher lobes of concentration kappa (1 to 1e6)
 * around the view directions (equal weights); weights times p / (alpha p + (1 - alpha) q).
 * Applies to the events of the kinds in species_mask up to scattering order max_order, as
 * monki_hip_attach_proposal; replaces any attached proposal's guidance settings.
 * monki_hip_run_directed is monki_hip_run with it. */
int monki_hip_set_directed(void *handle, float alpha, float kappa, uint32_t species_mask,
                           uint32_t max_order, char *error, size_t error_size);
int monki_hip_run_directed(void *handle, const uint32_t *controls, double *out, char *error,
                           size_t error_size);

/* Steering options of the guided and detector-directed tracers (both single-scene and batch
 * sessions; a batch session traces its rounds in the mode last set by monki_hip_attach_proposal
 * or monki_hip_set_directed). No reference counterpart.
 *   lobe_species: tabulated species whose bound CDF shapes lobes around the view directions
 *     (scattering cosines drawn from that species' sampler, measured from a view), or -1 for
 *     vMF lobes of the kappa of monki_hip_set_directed;
 *   lobe_fraction: guided tracer only, the lobes' share of the proposal (the rest is the
 *     network's mixture); the detector-directed tracer uses the lobes alone;
 *   surface: nonzero steers Lambertian reflections too (density mu / pi against the lobes);
 *   split_above, roulette_below, roulette_weight: the weight window on the product g of a
 *     photon's steering factors (its weight over the plain tracer's). 0 turns it off; else a
 *     photon whose g exceeds split_above after a scattering continues as up to 4 copies of
 *     equal weight, and one whose g is below roulette_below before a scattering survives with
 *     probability g / roulette_weight, its g raised to roulette_weight. Unbiased; bounds the
 *     products of the steering factors without touching the physical ones (albedos).
 * Returns 0, or nonzero with error text for invalid values. */
int monki_hip_set_steering(void *handle, int lobe_species, float lobe_fraction, int surface, float split_above,
                           float roulette_below, float roulette_weight, char *error, size_t error_size);

/* monki_hip_trace for the guided tracer (float32), same record layout. */
int monki_hip_trace_guided(void *handle, const uint32_t *controls, double *out, float *records,
                           size_t record_count, char *error, size_t error_size);

/* monki_hip_run (plain float32 transport) that also logs every scattering event into the
 * proposal's samples: features, direction after the event, and as weight the detector weight
 * the photon contributed after the event. controls[0] must not exceed the reserved photons. */
int monki_hip_run_logged(void *handle, const uint32_t *controls, double *out, void *proposal,
                         char *error, size_t error_size);

#ifdef __cplusplus
}
#endif

#endif
