/*
 * sce_oracle — CPU restatement of SCEMonteCarlo.jl's device-Metropolis-sweep path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * (scemontecarlo.jl_amd/csrc) and the timed CPU baseline of bench.py.  It is
 * never linked into, loaded by, or called from the product path.  It shares no
 * source with the product: every function here is an independent restatement of
 * the reference's algorithm (file:line cited per function in sce_oracle.c).
 *
 * Conventions: site ids, pointers and rows are 1-based exactly as the
 * reference's TiledHamiltonian stores them (src/hamiltonian.jl:234-266);
 * `config` is 3 x n_sites (column s = spin of site s), `zrows` is nlm x n_sites
 * column-major (column s contiguous).
 */
#ifndef SCE_ORACLE_H
#define SCE_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- keyed RNG (src/gpu/philox.jl) ---- */
void ora_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void ora_philox_block(uint64_t seed, int32_t site, int32_t sweep, uint32_t slot,
                      uint32_t tag, uint32_t out[4]);
double ora_uniform(uint32_t hi, uint32_t lo);
void ora_normal2(const uint32_t blk[4], double out[2]);

/* ---- project deterministic math (Cody-Waite reduction + Taylor series) ---- */
double ora_exp(double x);
double ora_log(double x);
/* the surrogate's FP32 activation (the device's det_tanhf) */
float ora_test_tanhf(float z);
void ora_sincos(double x, double* s, double* c);

/* ---- tesseral row (src/gpu/zlm_device.jl) ---- */
void ora_zlm_row(const double u[3], int lmax, double* z);
void ora_grad_zlm_row(const double u[3], int lmax, double* grow);

/* ---- tiled Hamiltonian (src/hamiltonian.jl) ---- */
typedef struct ora_tiled ora_tiled;
enum {
  ORA_INST_TERM = 0, ORA_INST_PTR, ORA_INST_SITES, ORA_SITE_PTR, ORA_SITE_INST,
  ORA_SITE_SLOT, ORA_SITE_HAS_L1, ORA_SITE_ACTIVE, ORA_SITE_PROG, ORA_SPROG_PTR,
  ORA_SENT_W, ORA_SENT_TGT, ORA_SFAC_PTR, ORA_SFAC_ROW, ORA_SFAC_SLOT,
  ORA_SITE_COL, ORA_SITE_COL2, ORA_PENT_ROW, ORA_PENT_ROW2, ORA_EPROG_PTR,
  ORA_EENT_W, ORA_EFAC_PTR, ORA_EFAC_ROW, ORA_TERM_COEF, ORA_COLOR_PTR,
  ORA_COLOR_SITES, ORA_NFIELDS
};
/* scalars: [n_cell_atoms, d1, d2, d3, n_sites, lmax, nlm, n_active, n_colors,
 *           n_terms, n_fitted_terms, n_inst] */
int ora_tiled_build(int n_cell_atoms, int n_terms, const int32_t* body,
                    const double* coef, const int32_t* atoms, const int32_t* shifts,
                    const int32_t* ls, const double* folded, const int32_t dims[3],
                    const double* magmoms, const double* field, ora_tiled** out,
                    char* err, size_t errlen);
void ora_tiled_free(ora_tiled* t);
void ora_tiled_scalars(const ora_tiled* t, int64_t out[12]);
const void* ora_tiled_array(const ora_tiled* t, int field, int64_t* len);

/* ---- energy contract (src/energy.jl) ---- */
void ora_zrows(const ora_tiled* t, const double* config, double* zrows);
double ora_total_energy(const ora_tiled* t, const double* zrows);
double ora_total_energy_tree(const ora_tiled* t, const double* zrows);
void ora_site_coeffs(const ora_tiled* t, int s, const double* zrows, double* c);
void ora_energy_gradient(const ora_tiled* t, const double* config, const double* zrows, int ws, double* G);
double ora_delta_energy(int nlm, const double* c, const double* zold, const double* znew);
double ora_reduce_tree(int64_t n, const double* x);

/* ---- keyed proposal + device walk contract ---- */
void ora_keyed_proposal(uint64_t seed, int32_t site, int32_t sweep, const double e[3],
                        double step, double e2[3], double* u_acc);
double ora_walk_delta(const ora_tiled* t, int s, const double* zrows,
                      const double* znew, int lanes);
void ora_walk_coeffs(const ora_tiled* t, int s, const double* zrows, int lanes,
                     double* c);
int64_t ora_keyed_sweep(const ora_tiled* t, double* config, double* zrows,
                        double* energy, double beta, double step, uint64_t seed,
                        int32_t sweep, int lanes, double* dE_q, int32_t* acc_q);
int64_t ora_keyed_or_sweep(const ora_tiled* t, double* config, double* zrows,
                           double* energy, double beta, uint64_t seed, int32_t sweep,
                           int32_t or_pass, int lanes, double* dE_q, int32_t* acc_q,
                           int64_t* natt);
double ora_renormalize(const ora_tiled* t, double* config, double* zrows, double* energy);

/* ---- learned delayed acceptance (an extension of the library) ---- */
int ora_sur_features(const ora_tiled* t, const double* config, int s, const double* e, float* out);
double ora_sur_delta(int F, const float* params, double scale, const float* phiA, const float* phiB);
double ora_sur_delta_q(int F, const float* qparams, double scale, const float* phiA, const float* phiB);
int64_t ora_da_sweep_q(const ora_tiled* t, double* config, double* zrows, double* energy, double beta,
                       double step, uint64_t seed, int32_t sweep, const float* const* qparams,
                       const double* scale, double* dE_q, int32_t* acc_q, int64_t* n_stage1);
int64_t ora_da_sweep(const ora_tiled* t, double* config, double* zrows, double* energy, double beta,
                     double step, uint64_t seed, int32_t sweep, const float* const* params,
                     const double* scale, double* dE_q, int32_t* acc_q, int64_t* n_stage1);

/* ---- reference CPU path (src/updates.jl, src/state.jl) ---- */
typedef struct ora_cpu_chain ora_cpu_chain;
ora_cpu_chain* ora_cpu_chain_create(const ora_tiled* t, const double* config,
                                    uint64_t seed, double step);
void ora_cpu_chain_free(ora_cpu_chain* c);
int64_t ora_cpu_metropolis_sweep(ora_cpu_chain* c, double beta, int nthreads);
int64_t ora_cpu_or_sweep(ora_cpu_chain* c, double beta, int nthreads);
double ora_cpu_renormalize(ora_cpu_chain* c);
void ora_cpu_state(const ora_cpu_chain* c, double* config, double* zrows,
                   double* energy, int64_t counters[4]);
double ora_cpu_energy(const ora_cpu_chain* c);
void ora_cpu_mean_spin(const ora_cpu_chain* c, double m[3]);
double ora_cpu_step(const ora_cpu_chain* c);
void ora_cpu_set_step(ora_cpu_chain* c, double step);
void ora_cpu_reset_counters(ora_cpu_chain* c);
void ora_cpu_counters(const ora_cpu_chain* c, int64_t counters[4]);

#ifdef __cplusplus
}
#endif
#endif
