/*
 * sce_hip.h — C ABI of libsce_hip.so, the MI355X (gfx950) device Metropolis
 * sweep of SCEMonteCarlo.jl.
 *
 * Every entry point returns SCE_OK (0) or a negative error code; the message of
 * the last failure on the calling thread is read with sce_last_error().  No C++
 * exception crosses this boundary.  Handles are not thread-safe; distinct handles
 * are.  Arrays follow the reference's Julia conventions unchanged: site ids,
 * CSR pointers and tesseral rows are 1-based Int32, `config` is 3 x n_sites and
 * `zrows` is nlm x n_sites, both column-major (column s contiguous).  The caller
 * owns every host buffer; the library copies in and out and owns device memory.
 *
 * Reference interface (Julia, src/SCEMonteCarlo.jl:86-88):
 *   GPUTiledHamiltonian(backend, H)          src/gpu/gpu_hamiltonian.jl:61-88
 *   GPUChainState(gH, st; seed)              src/gpu/gpu_state.jl:49-61
 *   gpu_metropolis_sweep!(gst, gH, β; ws)    src/gpu/gpu_sweep.jl:158-190
 *   gpu_run_sweeps!(gst, gH, st, β, n; ...)  src/gpu/gpu_sweep.jl:204-218
 *   to_host!(st, gst) / _from_host!(gst, st) src/gpu/gpu_state.jl:71-86
 *   _renormalize!(st, H)                     src/state.jl:168-185
 *   _swap_payload!(a, b)                     src/gpu/gpu_pt.jl:46-51
 *   gpu_run_pt(gH; kT, exchange_interval...) src/gpu/gpu_pt.jl:261-319
 */
#ifndef SCE_HIP_H
#define SCE_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SCE_OK = 0,
  SCE_EINVAL = -1,        /* ArgumentError in the reference */
  SCE_EDIM = -2,          /* DimensionMismatch */
  SCE_ERANGE_COUNTER = -3,/* 32-bit sweep counter exhausted (gpu_sweep.jl:163-164) */
  SCE_EHIP = -4,          /* HIP runtime failure */
  SCE_ENOMEM = -5,
  SCE_ELAYOUT = -6,       /* tables are not a translation-invariant tiling */
  SCE_ENODEV = -7         /* no HIP device */
};

typedef struct sce_tiled sce_tiled;   /* host TiledHamiltonian (tables owner) */
typedef struct sce_tables sce_tables; /* device tables (GPUTiledHamiltonian) */
typedef struct sce_chain sce_chain;   /* device chain (GPUChainState) */
typedef struct sce_batch sce_batch;   /* many chains in one launch (replica batch) */

/* The flat tables of one TiledHamiltonian (src/hamiltonian.jl:234-266,
 * _ContractionPrograms :62-91), as a Julia caller holds them. */
typedef struct sce_tables_desc {
  int32_t n_cell_atoms;
  int32_t dims[3];
  int32_t n_sites, lmax, nlm, n_terms, n_inst, n_active, n_colors, n_prog;
  int64_t n_inst_sites; /* = length(site_inst) = length(site_prog) = ... */
  int64_t n_sent, n_sfac, n_eent, n_efac;
  const int32_t *inst_term, *inst_ptr, *inst_sites;
  const int32_t *site_ptr, *site_inst;
  const int8_t *site_slot;
  const uint8_t *site_has_l1, *site_active;
  const int32_t *site_prog, *sprog_ptr;
  const double *sent_w;
  const int32_t *sent_tgt, *sfac_ptr, *sfac_row;
  const int8_t *sfac_slot;
  const int32_t *site_col, *site_col2, *pent_row, *pent_row2;
  const int32_t *eprog_ptr;
  const double *eent_w;
  const int32_t *efac_ptr, *efac_row;
  const double *term_coef;
  const int32_t *color_ptr, *color_sites;
} sce_tables_desc;

typedef struct sce_chain_info {
  double energy;        /* incremental energy, model units, intercept excluded */
  double step;          /* Metropolis rotation scale (radians) */
  double max_drift;     /* worst |E_inc - E_full| seen at renormalisation */
  double last_drift;
  int64_t sweep_index;  /* completed Metropolis sweeps: the keyed-RNG counter */
  int64_t acc_metro, att_metro, acc_or, att_or;
  uint64_t seed;
  int64_t acc_stage1;   /* delayed acceptance: proposals that passed the surrogate stage */
} sce_chain_info;

int32_t sce_version(void);
int32_t sce_last_error(char* buf, size_t len);
int32_t sce_device_count(int32_t* n);

/* ---- host TiledHamiltonian -------------------------------------------------
 * TiledHamiltonian(n_cell_atoms, terms; dims, magmoms, field)
 * (src/hamiltonian.jl:267-393).  Terms are flattened: term k has body[k] members;
 * atoms/ls are concatenated (1-based atoms), shifts are 3 ints per member,
 * folded is the column-major tensor of each term concatenated.  magmoms
 * (n_cell_atoms) and field (3) may be NULL. */
int32_t sce_tiled_build(int32_t n_cell_atoms, int32_t n_terms, const int32_t* body,
                        const double* coef, const int32_t* atoms, const int32_t* shifts,
                        const int32_t* ls, const double* folded, const int32_t dims[3],
                        const double* magmoms, const double* field, sce_tiled** out);
int32_t sce_tiled_desc(const sce_tiled* t, sce_tables_desc* out);
int32_t sce_tiled_free(sce_tiled* t);
/* Host-side rows and total energy: _zrows / _total_energy (src/energy.jl:41-75). */
int32_t sce_host_zrows(int32_t lmax, int64_t n, const double* config, double* zrows);
int32_t sce_host_total_energy(const sce_tables_desc* d, const double* zrows, double* E);

/* ---- device tables: GPUTiledHamiltonian(backend, H) -------------------------
 * Uploads the tables and derives the per-sublattice walk layout; rejects
 * lmax > 6 (gpu_hamiltonian.jl:67-68) and bodies > 4 with SCE_EINVAL. */
int32_t sce_tables_create(const sce_tables_desc* d, sce_tables** out);
int32_t sce_tables_destroy(sce_tables* t);
/* Derive and verify the walk layout on the host only (no device needed);
 * writes the same facts as sce_tables_layout_info. */
int32_t sce_layout_check(const sce_tables_desc* d, int64_t out[5]);
/* Layout facts: [n_sublattices, max_neighbours, max_walk_entries,
 * total_walk_entries, fmax = widest surrogate descriptor row] */
int32_t sce_tables_layout_info(const sce_tables* t, int64_t out[5]);
/* The per-position tables sce_tables_create builds on the device (no reference
 * counterpart; diagnostics): for every colour position q (n_active of them),
 * the nb_max neighbour sites (0-based, padded with the site itself), the
 * lower-colour mask of those slots (when nb_max <= 64) and the raw adjacency
 * index -> pid-sorted index permutation (adj_max per position).  sizes =
 * [nb_max, adj_max, has_mask]; any array may be NULL (sizes only).
 * sce_position_tables_host makes the same tables on the host (no device). */
int32_t sce_tables_position_tables(sce_tables* t, int32_t* nbq, uint64_t* low, uint16_t* perm,
                                   int64_t sizes[3]);
int32_t sce_position_tables_host(const sce_tables_desc* d, int32_t* nbq, uint64_t* low, uint16_t* perm,
                                 int64_t sizes[3]);

/* ---- device chain: GPUChainState(gH, st; seed) ------------------------------ */
int32_t sce_chain_create(sce_tables* t, const double* config, const double* zrows,
                         double energy, uint64_t seed, double step, sce_chain** out);
int32_t sce_chain_destroy(sce_chain* c);
int32_t sce_chain_info_get(sce_chain* c, sce_chain_info* out);
int32_t sce_chain_set_step(sce_chain* c, double step);
int32_t sce_chain_reset_counters(sce_chain* c);
/* Restores the lane-owned counters of a chain (gpu_state.jl:19-47: sweep_index
 * keys the next sweep's Philox counter as sweep_index + 1) — with sce_upload,
 * what resuming a checkpointed chain needs. */
int32_t sce_chain_set_counters(sce_chain* c, int64_t sweep_index, int64_t acc_metro,
                               int64_t att_metro, int64_t acc_or, int64_t att_or);
/* Restores the chain's drift record (the largest |incremental - recomputed|
 * energy seen at a renormalisation), which a checkpoint carries with the
 * chain state (max_drift, src/checkpoint.jl:165-190). */
int32_t sce_chain_set_drift(sce_chain* c, double max_drift);
/* to_host!(st, gst): config (3 x n), zrows (nlm x n), energy; any may be NULL. */
int32_t sce_download(sce_chain* c, double* config, double* zrows, double* energy);
/* _from_host!(gst, st) */
int32_t sce_upload(sce_chain* c, const double* config, const double* zrows, double energy);
/* _swap_payload!(a, b): exchanges config / zrows / energy (O(1)). */
int32_t sce_swap_payload(sce_chain* a, sce_chain* b);

/* gpu_metropolis_sweep!(gst, gH, β; workgroupsize): one Metropolis sweep;
 * `lanes` (16, 32 or 64) is the number of lanes that walk one site — part of
 * the determinism scope, as the reference's workgroupsize.  128 and 256 (two
 * and four wavefronts per site, the fold's cross-wavefront steps through LDS)
 * are accepted here and by sce_run_sweeps for Metropolis-only sweeps with
 * per-colour launches.
 * Writes the number of accepted moves. */
int32_t sce_metropolis_sweep(sce_chain* c, double beta, int32_t lanes, int64_t* nacc);
/* One overrelaxation sweep (updates.jl:66-87,231-290 on the device); `or_pass`
 * selects the keyed accept stream of the pass within a compound sweep. */
int32_t sce_overrelaxation_sweep(sce_chain* c, double beta, int32_t or_pass,
                                 int32_t lanes, int64_t* nacc);
/* gpu_run_sweeps!(gst, gH, st, β, n; renorm_interval, workgroupsize) without the
 * final to_host!: n compound sweeps (1 Metropolis + or_per_metropolis OR each),
 * renormalising on the device every renorm_interval sweeps (<= 0: never).
 * When measure_interval > 0, every measure_interval-th sweep appends
 * (E, m_x, m_y, m_z) to `obs` (capacity obs_cap rows), m = mean active spin. */
int32_t sce_run_sweeps(sce_chain* c, double beta, int64_t nsweeps,
                       int32_t or_per_metropolis, int64_t renorm_interval,
                       int64_t measure_interval, int32_t lanes, double* obs,
                       int64_t obs_cap, int64_t* n_obs);
/* Run the Metropolis-only sweeps of sce_run_sweeps as one persistent
 * cooperative launch per renormalisation segment (on != 0) or as per-colour
 * launches (0, the default).  Both give bit-identical results. */
int32_t sce_chain_set_persistent(sce_chain* c, int32_t on);
/* Replay each sweep of sce_run_sweeps as a captured hipGraph (on != 0, the
 * default) or launch its kernels directly; results are identical. */
int32_t sce_chain_set_graph(sce_chain* c, int32_t on);
/* Streaming error analysis on the device, fed by every measurement row of
 * later sce_run_sweeps / sce_da_run_sweeps calls: the LogBinner of
 * src/binning.jl:25-79 and the BinStore of :138-188 over the five series
 * E, E^2, |m|, m^2, m^4 (m^2 = (m_x^2 + m_y^2) + m_z^2).  reset zeroes the
 * state and turns it on (bin_size >= 1 and 2 <= nbins <= 1024, or both 0 for
 * the LogBinner alone); off turns it off. */
int32_t sce_chain_binner_reset(sce_chain* c, int32_t bin_size, int32_t nbins);
int32_t sce_chain_binner_off(sce_chain* c);
/* The binner state: n pushes; count[62]; sums[62][5], sums2[62][5] (row
 * major); nfull completed BinStore bins and their means[nfull][5] (means_cap
 * rows available). */
int32_t sce_chain_binner_get(sce_chain* c, int64_t* n, int64_t* count, double* sums, double* sums2,
                             int32_t* nfull, double* means, int32_t means_cap);
/* _renormalize!(st, H) on the device: normalise active spins, rebuild rows,
 * recompute the FP64 total energy, re-anchor.  Writes the drift. */
int32_t sce_renormalize(sce_chain* c, double* drift);
/* Device total energy of the chain's current rows (no re-anchoring). */
int32_t sce_chain_total_energy(sce_chain* c, double* E);
/* Enable (on != 0) HIP-event timing of sce_run_sweeps: the wall time of the
 * device work and the per-launch durations of the Metropolis walk kernel. */
int32_t sce_chain_set_timing(sce_chain* c, int32_t on);
/* Wall time of the device work enqueued by the last timed sce_run_sweeps call, and
 * the average duration of its Metropolis walk launches (both from HIP events). */
int32_t sce_chain_timing(sce_chain* c, double* total_ms, double* walk_ms_avg,
                         int64_t* walk_launches);

/* Average duration of the colour-walk kernel over `reps` back-to-back passes of
 * every colour class between one pair of HIP events (a diagnostic with no
 * reference counterpart; the chain's config, rows, energy and counters are left
 * as they were). */
int32_t sce_chain_time_walk(sce_chain* c, int32_t lanes, int32_t reps, double* ms_avg,
                            int64_t* launches);

/* ---- batches: many chains in one launch per colour --------------------------
 * n_chains chains, each on its own tables (a replica ladder shares one tables
 * handle; uneven batches mix models and sizes with equal lmax), with its own
 * keyed stream seed, step and kT.  Every chain is bit-identical to the same
 * chain run alone.  With exchange_interval > 0 (all chains on the same tables,
 * kT ladder in chain order) adjacent chains exchange payloads on the device
 * with the rule of _swap_accepts (src/pt.jl:92-96) and the alternating-parity
 * schedule of _gpu_pt_phase! (src/gpu/gpu_pt.jl:115-140); the exchange uniform
 * is keyed by (exchange_seed, pair, exchange count). */
int32_t sce_batch_create(int32_t n_chains, sce_tables* const* tables,
                         const double* const* configs, const double* const* zrows,
                         const double* energies, const uint64_t* seeds, const double* steps,
                         const double* kts, uint64_t exchange_seed, sce_batch** out);
int32_t sce_batch_destroy(sce_batch* b);
/* nsweeps sweeps of every chain.  Sweep j = phase_done + i (i = 1..nsweeps) of
 * the current phase renormalises when j % renorm_interval == 0, measures
 * (E, m) of every chain into obs (chain-major, obs_cap rows per chain) when
 * j % measure_interval == 0, and exchanges when j % exchange_interval == 0 and
 * j < phase_total (no exchange after a phase's last sweep, gpu_pt.jl:128). */
int32_t sce_batch_run(sce_batch* b, int64_t nsweeps, int32_t lanes, int64_t renorm_interval,
                      int64_t measure_interval, int64_t exchange_interval, int64_t phase_done,
                      int64_t phase_total, double* obs, int64_t obs_cap, int64_t* n_obs);
int32_t sce_batch_download(sce_batch* b, int32_t k, double* config, double* zrows,
                           double* energy);
int32_t sce_batch_info(sce_batch* b, int32_t k, sce_chain_info* out);
int32_t sce_batch_set_step(sce_batch* b, int32_t k, double step);
/* Zeroes every chain's acceptance counters and drift record (exchange
 * statistics accumulate over the whole run, as in the reference). */
int32_t sce_batch_reset_counters(sce_batch* b);
/* _renormalize! of every chain on the device (state.jl:168-185). */
int32_t sce_batch_renormalize(sce_batch* b);
/* Attempted / accepted exchanges of pair (k, k+1), k = 0..n_chains-2. */
int32_t sce_batch_swap_stats(sce_batch* b, int64_t* att, int64_t* acc);
/* Replay captured sweeps as hipGraphs (on != 0, default) or launch directly. */
int32_t sce_batch_set_graph(sce_batch* b, int32_t on);
/* Run the sweeps of sce_batch_run, replica exchange included, as one
 * persistent cooperative launch for every chain per renormalisation segment
 * (on != 0; stencils of at most 64 neighbour slots, (lmax + 1)^2 <= lanes) or
 * as per-colour launches of every chain (0, the default).  Both give
 * bit-identical results at the same lanes per site.  The reference runs rungs round-robin on one queue
 * (src/gpu/gpu_pt.jl:4-15). */
int32_t sce_batch_set_persistent(sce_batch* b, int32_t on);
/* At 64 lanes per site, one wavefront walks a site's sublattice walk once for
 * `chains` (2 or 4) chains of a batch whose chains all share one tables
 * handle (a temperature ladder), each chain's sums formed as at 64 lanes per
 * site alone: the results are those of 64 lanes per site, each walk entry is
 * loaded once per `chains` sites.  0 (the default): 2 for such batches of at
 * least two chains whose neighbour stencil fits, else off; 1 turns it off;
 * batches of different tables return SCE_EINVAL for 2 and 4.  Per-colour
 * launches only.
 * The reference has no counterpart: its rungs run one after another on one
 * queue (src/gpu/gpu_pt.jl:4-15). */
int32_t sce_batch_set_shared_walk(sce_batch* b, int32_t chains);
/* At 64 lanes per site, one wavefront walks two sites of one chain's colour
 * class on the same sublattice (on != 0), for batches whose chains do not
 * share walks across chains (sce_batch_set_shared_walk off or inapplicable):
 * each entry of the sublattice walk is loaded once for both sites; the results
 * are those of 64 lanes per site.  Per-colour launches only; off by default.
 * The reference has no counterpart (its device kernel takes one site per
 * work-item group, src/gpu/gpu_sweep.jl). */
int32_t sce_batch_set_site_pairs(sce_batch* b, int32_t on);
/* Checkpoint restore of a "gpu_pt" run (resume, src/checkpoint.jl:555-635).
 * set_counters: the lane-owned state of chain k that batch creation does not
 * take — its keyed-RNG sweep index, acceptance counters and drift record
 * (gst.sweep_index / acc_metro / att_metro and st.max_drift restored at
 * src/checkpoint.jl:584-597).  get/set_exchange: the exchange schedule's state
 * — exchanges attempted so far (the keyed exchange counter), the parity of the
 * next attempt, and the per-pair statistics (n_chains - 1 each); the
 * reference stores its exchange Xoshiro words, parity and swap_att/swap_acc
 * (src/checkpoint.jl:396-400). */
int32_t sce_batch_set_counters(sce_batch* b, int32_t k, int64_t sweep_index, int64_t acc_metro,
                               int64_t att_metro, int64_t acc_or, int64_t att_or, double max_drift);
int32_t sce_batch_get_exchange(sce_batch* b, int64_t* exch_count, int32_t* parity);
int32_t sce_batch_set_exchange(sce_batch* b, int64_t exch_count, int32_t parity, const int64_t* att,
                               const int64_t* acc);

/* ---- learned delayed acceptance (no reference counterpart: an extension of the
 * device sweep for callers that want more effective samples per second) ------
 * A surrogate holds, per sublattice of the tables, a small MLP g(phi(e)) of a
 * rotation-invariant descriptor phi of a site's spin e against its neighbour
 * spins (features: e.e_n and (e.e_n)^2 per stencil neighbour, e.(e_x x e_y) per
 * triplet, the three pairings of each body-4 cluster, the spin components when
 * a body-1 term touches the site).  Network: W1 (64 x F) tanh, W2 (64 x 64)
 * tanh, w3 (64) plus a linear path v (F); parameter vector
 * [W1 row-major, b1, W2 row-major, b2, w3, v] (FP32), output scale (FP64).
 * A delayed-acceptance sweep (Christen & Fox 2005) first accepts the keyed
 * proposal with min(1, exp(-beta dE_s)), dE_s = scale (g(phi(e')) - g(phi(e))),
 * and only then walks the exact FP64 dE and accepts with
 * min(1, exp(-beta (dE - dE_s))) (keyed uniform: slot 3).  Detailed balance
 * holds for any surrogate; the exact dE decides. */
typedef struct sce_surrogate sce_surrogate;
/* Glorot-initialised networks (splitmix64 stream from seed), v = 0, scale 1. */
int32_t sce_sur_create(sce_tables* t, uint64_t seed, sce_surrogate** out);
int32_t sce_sur_destroy(sce_surrogate* s);
/* Features F and parameter count of sublattice `sub` (F = 0: inactive
 * sublattice, no network), its output scale, and the harvested row count. */
int32_t sce_sur_info(const sce_surrogate* s, int32_t sub, int32_t* F, int64_t* n_params,
                     double* scale, int64_t* n_rows);
int32_t sce_sur_get_params(sce_surrogate* s, int32_t sub, float* out);
int32_t sce_sur_set_params(sce_surrogate* s, int32_t sub, const float* in, double scale);
/* Stage 1 of later delayed-acceptance sweeps on the int8 networks (on != 0:
 * quantised descriptor and weights, exact integer dot products on the matrix
 * cores) or on the FP32 networks (0, the default).  Either way the exact FP64
 * dE decides.  No reference counterpart (the surrogate is new). */
int32_t sce_sur_set_int8(sce_surrogate* s, int32_t on);
/* The int8 network of sublattice `sub` as floats: V1 (64 x F), r1 (64), b1 (64),
 * V2 (64 x 64), r2 (64), b2 (64), w3 (64), v (F), input scales (F); refreshed
 * whenever the FP32 parameters change (set_params, training). */
int32_t sce_sur_get_qparams(sce_surrogate* s, int32_t sub, float* out);
/* nsweeps plain Metropolis sweeps of chain c (the same chain evolution as
 * sce_run_sweeps) that also store (phi(e), phi(e'), exact dE) of every attempt
 * as training rows on the device. */
int32_t sce_sur_harvest(sce_surrogate* s, sce_chain* c, double beta, int64_t nsweeps);
/* Adam over sublattice sub's rows: `epochs` shuffled passes in minibatches of
 * bsz pairs, loss mean((scale (g_B - g_A) - dE)^2) / scale^2 (scale fixed to
 * the rms dE of the rows at the first call); per-epoch mean loss in loss_out. */
int32_t sce_sur_train(sce_surrogate* s, int32_t sub, int32_t epochs, int32_t bsz, double lr,
                      uint64_t seed, double* loss_out);
/* Loss and its gradient (layout of the parameter vector) of n given pairs
 * through the training kernels; A, B: n x fmax FP32 descriptor rows. */
int32_t sce_sur_grad(sce_surrogate* s, int32_t sub, int32_t n, const float* A, const float* B,
                     const double* y, float* grad_out, double* loss);
/* dE_s of n given descriptor pairs through the sweep's inference path. */
int32_t sce_sur_eval(sce_surrogate* s, int32_t sub, int32_t n, const float* A, const float* B,
                     double* out);
/* Descriptor rows phi(e_site), phi(e2) (n x fmax FP32, zero padded) of given
 * active sites (1-based) and proposed spins e2 (3 x n) against config. */
int32_t sce_sur_features(sce_tables* t, const double* config, int32_t n, const int32_t* sites,
                         const double* e2, float* outA, float* outB);
/* nsweeps delayed-acceptance sweeps of chain c with the same renormalisation
 * and measurement schedule as sce_run_sweeps (hipGraph replay unless
 * sce_chain_set_graph(c, 0)); acc_stage1 in sce_chain_info counts the proposals
 * that passed the surrogate stage. */
int32_t sce_da_run_sweeps(sce_chain* c, sce_surrogate* s, double beta, int64_t nsweeps,
                          int64_t renorm_interval, int64_t measure_interval, double* obs,
                          int64_t obs_cap, int64_t* n_obs);

/* ---- all-site energy gradient ----------------------------------------------
 * gpu_energy_gradient!(dG, gH, dconfig, gsc; workgroupsize = lanes,
 * refresh_zrows = true) — src/gpu/gpu_gradient.jl:187-212: the tangent-projected
 * dE/de_s of every site (3 x n_sites, model units, intercept excluded; inactive
 * sites exactly zero) for the configuration `config` (3 x n_sites, unit spins),
 * whose rows are rebuilt on the device.  The arithmetic is the reference's lane
 * contract (_gradient_lane_ref!, :242-279) with `lanes` strided shares
 * (64, 128 or 256; the reference's default is 128), so the result is the same
 * bits on any backend. */
int32_t sce_energy_gradient(sce_tables* t, const double* config, double* G, int32_t lanes);
/* No counterpart in the reference: sce_energy_gradient for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and bits; the device copy of
 * config, its rebuilt rows and the gradient live in one static device array of
 * 4 Mi doubles (32 MiB) per GPU, so (3 + nlm + 3) * n_sites must not exceed
 * 4 Mi (SCE_EDIM otherwise).  It calls no hipMalloc/hipFree.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU race. */
int32_t sce_energy_gradient_noalloc(sce_tables* t, const double* config, double* G, int32_t lanes);
/* gpu_energy_gradient!(dG, gst, gH, gsc) — :222-226: the same on the chain's
 * device spins and rows (no rebuild: the rows are Z(config) by the sweep
 * invariant). */
int32_t sce_chain_energy_gradient(sce_chain* c, double* G, int32_t lanes);
/* Average gradient-kernel duration over reps back-to-back launches on the
 * chain's rows (measurement; no reference counterpart). */
int32_t sce_chain_time_gradient(sce_chain* c, int32_t lanes, int32_t reps, double* ms_avg);

/* ---- diagnostics (no reference counterpart): the device arithmetic exposed
 * for bit-exact parity tests ------------------------------------------------- */
int32_t sce_test_philox(int64_t n, const uint64_t* seeds, const uint32_t* ctr4,
                        uint32_t* out4);
/* The placement of a chained sweep's proposals (host only, no device): for
 * classes with 0-based offsets color_ptr[0..n_colors], records (launch, lo, n,
 * sweep offset) of the proposal ranges; launch -1 = before the walks, 0..K-1 =
 * the K non-empty classes' walk launches, K = the reduction launch, K + 1 =
 * after it.  kf: the walk launch carrying a pending final (-1 if none). */
int32_t sce_test_chain_plan(int32_t n_colors, const int32_t* color_ptr, int32_t next_prop, int32_t pending,
                            int32_t carry_last, int32_t walk_props, int32_t* out, int32_t cap, int32_t* n_out,
                            int32_t* kf);
/* op: 0 exp, 1 log, 2 sin, 3 cos, 4 the surrogate's FP32 tanh (of x rounded to float) */
int32_t sce_test_detmath(int64_t n, int32_t op, const double* x, double* out);
int32_t sce_test_zlm(int64_t n, int32_t lmax, const double* dirs3, double* out);
int32_t sce_test_proposal(int64_t n, uint64_t seed, const int32_t* sites,
                          const int32_t* sweeps, const double* e3, double step,
                          double* e2_out, double* u_out);
/* ΔE of the kernel walk for arbitrary proposals: sites (1-based, active),
 * znew (nlm x n), against rows `zrows` (nlm x n_sites). */
int32_t sce_test_walk_delta(sce_tables* t, const double* zrows, int64_t n,
                            const int32_t* sites, const double* znew, int32_t lanes,
                            double* out);
/* Leave-one-out coefficients c (nlm x n) of the overrelaxation walk. */
int32_t sce_test_walk_coeffs(sce_tables* t, const double* zrows, int64_t n,
                             const int32_t* sites, int32_t lanes, double* out);
/* Device deterministic tree sum of x (the sweep/energy reduction). */
int32_t sce_test_reduce(int64_t n, const double* x, double* out);
/* One v_mfma_f32_16x16x4f32 per test t: out[t] = c[t] + sum_k a[4t+k] b[4t+k] as
 * the matrix core computes it (the accumulation order the oracle restates). */
int32_t sce_test_mfma(int64_t n, const float* a4, const float* b4, const float* c, float* out);

#ifdef __cplusplus
}
#endif
#endif
