// sce_hip.hip — gfx950 kernels and the C ABI of the device Metropolis sweep.
//
// Reference path: src/gpu/gpu_sweep.jl (the fused colour kernel, :92-145, and
// its drivers, :158-218), src/gpu/gpu_state.jl, src/gpu/gpu_hamiltonian.jl,
// src/updates.jl (overrelaxation, :66-87), src/state.jl (_renormalize!,
// :168-185), src/energy.jl (_total_energy, :55-75).
//
// One Metropolis sweep is
//   k_propose            four lanes per active site: keyed proposal + Z row
//   k_metro   x colours  `lanes` lanes per site: neighbour rows staged in LDS,
//                        the site's sublattice walk dealt round-robin over the
//                        lanes, fma-accumulated ΔE, shuffle-tree fold, FP64
//                        accept; per-site writes only (no atomics on state)
//   k_chunks + k_final   deterministic tree sum of the staged ΔE + counters
//                        (+ the mean spin when measuring), folded into the
//                        device-resident chain state
// all on the chain's stream with no host synchronisation between sweeps; a
// sweep is captured once as a hipGraph and replayed.  Within one call the
// sweeps are chained (enqueue_metro): the next sweep's proposals ride in the
// walk launches and the final reduction in the next sweep's smallest-class
// walk, so a chained sweep is the colour walks plus one reduction launch.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <queue>
#include <vector>

#include "sce_internal.h"


// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[1024] = "";

int32_t sce_fail(int32_t code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      return sce_fail(SCE_EHIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                      __LINE__);                                                     \
  } while (0)

#define CHK(x)                  \
  do {                          \
    int32_t r_ = (x);           \
    if (r_ != SCE_OK) return r_; \
  } while (0)

extern "C" int32_t sce_version(void) { return 10000; }

extern "C" int32_t sce_last_error(char* buf, size_t len) {
  if (buf && len) {
    std::strncpy(buf, g_err, len - 1);
    buf[len - 1] = 0;
  }
  return SCE_OK;
}

extern "C" int32_t sce_device_count(int32_t* n) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  if (n) *n = c;
  return c > 0 ? SCE_OK : sce_fail(SCE_ENODEV, "no HIP device available");
}

// ---------------------------------------------------------------------------
// device-side structures
// ---------------------------------------------------------------------------
struct ChainDev {
  double energy, step, beta, max_drift, last_drift, e_full;
  long long sweep_index, acc_metro, att_metro, acc_or, att_or, nacc_last, acc_stage1;
  unsigned long long seed;
  unsigned long long iacc, iatt;
  unsigned int ticket;
  int n_obs;
};

struct DevTables {
  int n_sites, n_active, nlm, lmax, na, d0, d1, d2, nb_max, n_inst;
  const int32_t* color_sites;  // 0-based site ids, colour order
  const uint8_t* has_l1;
  const uint8_t* active;
  const int32_t* walk_ptr;
  const WalkEntry* walk;
  // the same walks, each padded with no-op entries to a multiple of WALK_PAD
  // (what walk_delta reads: no bounds test or clamp per entry)
  const int32_t* walkp_ptr;
  const WalkEntry* walkp;
  // per sublattice: the leading entries of its padded walk with one factor
  // (the others on the 1.0 slot), and with at most two
  const int2* walkp_pre;
  const int32_t* nb_ptr;
  const int4* nb;
  const int32_t* nbq;  // per colour position: nb_max neighbour site ids (padded with the site's own)
  const unsigned long long* nbq_low;  // per position: bit j = slot j is an active site of lower colour (nb_max <= 64)
  const int32_t* prog_ptr;  // na + 1: per sublattice, into prog_off
  const int32_t* prog_off;  // walk offsets of the pid-sorted adjacency's programs (+ end)
  const uint16_t* adj_perm; // per colour position: raw adjacency index -> pid-sorted index
  int adj_max;              // adjacency entries per site, at most (the adj_perm stride)
  const int32_t* inst_term;  // 0-based
  const int32_t* inst_ptr;   // 0-based offsets
  const int32_t* inst_sites; // 0-based
  const int32_t* eprog_ptr;  // 0-based offsets
  const double* eent_w;
  const int32_t* efac_ptr;   // 0-based offsets
  const int32_t* efac_row;   // 0-based rows
  const double* term_coef;
  const int32_t* feat_ptr;   // surrogate descriptor specs per sublattice
  const int4* feat;
  int nfeat_max;
};

// The reference's adjacency programs, 1-based as the TiledHamiltonian holds
// them (src/hamiltonian.jl site programs), for the all-site energy gradient,
// whose arithmetic contract deals whole programs to lanes.
struct GradTables {
  int n_sites, nlm;
  const int32_t *site_ptr, *site_inst, *site_prog, *site_col, *site_col2;
  const int32_t *sprog_ptr, *sent_tgt, *sfac_ptr, *sfac_row, *pent_row, *pent_row2;
  const int32_t *inst_ptr, *inst_sites;
  const int8_t* sfac_slot;
  const double* sent_w;
};

constexpr int TPB = 256;      // threads per block of the walk kernels

// int8 stage-1 inference: descriptor features are quantised as rint(x * s)
// with s = 127 for features bounded by 1 and 42 for the four-body pairing sums
// (bounded by 3); each layer's weights per output unit j as
// rint(w / s_in * 127 / t_j), t_j the largest |w / s_in| of the unit.
constexpr float SUR_QIN_UNIT = 127.0f, SUR_QIN_PAIRS = 42.0f;
__host__ __device__ inline int sur_fpq(int F) { return (F + 63) / 64 * 64; }
constexpr int TREE_B = 512;   // elements per tree chunk

// A load from global memory through a pointer the compiler cannot place: the
// batch kernels read their chains' pointers from device memory, and loads
// through those generic pointers would compile to flat instructions, which
// also count against the LDS counter (a wait on an LDS gather would wait on
// them too).
template <class T>
__device__ __forceinline__ T ldg(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ int2 ldg_int2(const int2* p) {
  const long long v = ldg(reinterpret_cast<const long long*>(p));
  return make_int2((int)v, (int)(v >> 32));
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each XCD
// walks a contiguous range of the colour class (neighbour rows then stay in
// that XCD's L2).
__device__ __forceinline__ int xcd_swizzle(int b, int nb) {
  const int x = b & 7, l = b >> 3, per = nb >> 3, rem = nb & 7;
  return x * per + (x < rem ? x : rem) + l;
}

// ---------------------------------------------------------------------------
// proposals at sweep start (a site's spin and row change only at its own
// colour step, so this equals proposing there): four lanes per active site.
// Lanes 4i .. 4i+3 evaluate the radius and the angle of both Box-Muller pairs
// of sce::keyed_proposal side by side (philox slots 1 and 2), exchange them,
// and lane 4i completes the proposal and its Z row.  The operations are those
// of sce::keyed_proposal / normal2, so the bits are the same.
// ---------------------------------------------------------------------------
constexpr int PROPOSE_LANES = 4;

// Thread t of the proposal pass (PROPOSE_LANES per site) for the sweep whose
// 1-based index is the chain's completed sweeps + sweep_off.
// The keyed proposal of site s (0-based) for the 1-based sweep `sweep` from
// spin e, by a group of PROPOSE_LANES lanes (part = lane within the group; all
// four take part in the exchange): each lane draws its half of a Box-Muller
// pair, the radius (even part) or the angle.  Valid in part 0 only (returns
// true there): e2, the accept uniform u and the row z of e2.
template <int LMAX>
__device__ __forceinline__ bool propose_core(uint64_t seed, int s, int32_t sweep, const double e[3], double step,
                                             int part, double e2[3], double* u_out,
                                             double z[(LMAX + 1) * (LMAX + 1)]) {
  const sce::u32x4 b = sce::philox_block(seed, s + 1, sweep, 1u + (uint32_t)(part >> 1), 0u);
  double x0 = 0.0, x1 = 0.0;
  if ((part & 1) == 0)
    x0 = __builtin_sqrt(-2.0 * sce::det_log(sce::uniform01(b.x, b.y)));
  else
    sce::det_sincos(6.283185307179586 * sce::uniform01(b.z, b.w), &x0, &x1);
  const int l0 = (int)(threadIdx.x % 64) & ~(PROPOSE_LANES - 1);
  const double r1 = __shfl(x0, l0), s1 = __shfl(x0, l0 + 1), c1 = __shfl(x1, l0 + 1);
  const double r2 = __shfl(x0, l0 + 2), s2 = __shfl(x0, l0 + 3), c2 = __shfl(x1, l0 + 3);
  if (part != 0) return false;
  const sce::u32x4 b0 = sce::philox_block(seed, s + 1, sweep, 0u, 0u);
  const double u_flip = sce::uniform01(b0.x, b0.y);
  *u_out = sce::uniform01(b0.z, b0.w);
  if (u_flip < 0.2) {
    e2[0] = -e[0];
    e2[1] = -e[1];
    e2[2] = -e[2];
  } else {
    const double n1 = r1 * c1, n2 = r1 * s1, n3 = r2 * c2, n4 = r2 * s2;
    const double nn = __builtin_sqrt((n1 * n1 + n2 * n2) + n3 * n3);
    const double ax[3] = {n1 / nn, n2 / nn, n3 / nn};
    sce::rotate3(e, ax, step * n4, e2);
  }
  sce::zlm_row<LMAX>(e2[0], e2[1], e2[2], z);
  return true;
}

template <int LMAX>
__device__ __forceinline__ void propose_body(const DevTables& T, const double* __restrict__ config,
                                             double* __restrict__ prop, const ChainDev* __restrict__ st,
                                             int q_lo, int q_n, int t, int sweep_off) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1), PS = NLM + 4;
  const int qi = t / PROPOSE_LANES, part = t % PROPOSE_LANES;
  const bool valid = qi < q_n;
  const int q = q_lo + (valid ? qi : q_n - 1);  // every lane takes part in the exchange
  const int s = T.color_sites[q];
  const int32_t sweep = (int32_t)(st->sweep_index + sweep_off);
  const double e[3] = {config[3 * (size_t)s], config[3 * (size_t)s + 1], config[3 * (size_t)s + 2]};
  double e2[3], u, z[NLM];
  const bool mine = propose_core<LMAX>(st->seed, s, sweep, e, st->step, part, e2, &u, z);
  if (!valid || !mine) return;
  double* o = prop + (size_t)q * PS;
  o[0] = e2[0];
  o[1] = e2[1];
  o[2] = e2[2];
  o[3] = u;
#pragma unroll
  for (int k = 0; k < NLM; ++k) o[4 + k] = z[k];
}

template <int LMAX>
__global__ __launch_bounds__(256) void k_propose(DevTables T, const double* __restrict__ config,
                                                 double* __restrict__ prop,
                                                 const ChainDev* __restrict__ st, int q_lo,
                                                 int q_n, int sweep_off) {
  propose_body<LMAX>(T, config, prop, st, q_lo, q_n, blockIdx.x * blockDim.x + threadIdx.x, sweep_off);
}

// Proposals of the colour positions [lo, lo + n) for the sweep whose 1-based
// index is the completed sweeps + off, made as extra blocks of another launch.
struct PropRange {
  int lo, n, off;
};
__host__ __device__ inline int prop_blocks(int n) { return (int)(((long long)n * PROPOSE_LANES + 255) / 256); }

// Stage the neighbour rows of site s into Z (nb*NLM + row) and 1.0 at NB*NLM.
// Each lane fetches STAGE_ROUND(L) entries per round with every load issued
// before the first is used: indices past the end are clamped to the last entry
// (an in-bounds load whose value is written to that entry's own slot again), so
// no lane-dependent branch splits the round into serial load-use pairs.
__host__ __device__ constexpr int stage_round(int L) { return L >= 64 ? 4 : 8; }

template <int L, int NLM>
__device__ __forceinline__ void stage_neighbours(const DevTables& T, const double* __restrict__ zrows,
                                                 int s, int lane, double* Z, int* atom_out) {
  constexpr int U = stage_round(L);
  const int atom = s % T.na;
  const int cell = s / T.na;
  const int c1 = cell % T.d0;
  const int r = cell / T.d0;
  const int c2 = r % T.d1;
  const int c3 = r / T.d1;
  const int nb0 = T.nb_ptr[atom];
  const int total = (T.nb_ptr[atom + 1] - nb0) * NLM;
  for (int b = 0; b < total; b += U * L) {
    int at[U];
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = b + u * L + lane;
      at[u] = i < total ? i : total - 1;
      const int nbi = at[u] / NLM;
      const int row = at[u] - nbi * NLM;
      const int4 o = T.nb[nb0 + nbi];
      int x = c1 + o.y, y = c2 + o.z, z = c3 + o.w;
      if (x >= T.d0) x -= T.d0;
      if (y >= T.d1) y -= T.d1;
      if (z >= T.d2) z -= T.d2;
      const int n = o.x + T.na * (x + T.d0 * (y + T.d1 * z));
      v[u] = zrows[(size_t)n * NLM + row];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) Z[at[u]] = v[u];
  }
  if (lane == 0) Z[total] = 1.0;
  *atom_out = atom;
}

// The same staging for site s at colour position q, from the per-position
// neighbour list: every slot of nb_max rows is filled (padding repeats the
// site's own row, which no walk entry reads), so the row loads wait on one
// load, issued first, whose address is the position.  between() runs once,
// after the first list loads are issued (loads that need s go there).
// The 1.0 slot follows the sublattice's own rows, so for a sublattice with
// fewer than nb_max neighbours it lies inside the padding.  One wavefront
// stores it after its padding stores, in program order; a group of several
// wavefronts (L > 64) has no such order across them, so there the padding
// store of that slot writes 1.0 itself.
// Returns the sublattice.
template <int L, int NLM, class F>
__device__ __forceinline__ int stage_neighbours_q(const DevTables& T, const double* __restrict__ zrows,
                                                  int q, int s, int lane, double* Z, F&& between) {
  constexpr int U = stage_round(L);
  const int32_t* nq = T.nbq + (size_t)q * T.nb_max;
  const int total = T.nb_max * NLM;
  const int atom = s % T.na;
  int one = -1;  // the 1.0 slot, where the padding stores must not reach it last
  if constexpr (L > 64) one = (ldg(T.nb_ptr + atom + 1) - ldg(T.nb_ptr + atom)) * NLM;
  for (int b = 0; b < total; b += U * L) {
    int at[U], n[U];
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = b + u * L + lane;
      at[u] = i < total ? i : total - 1;
      n[u] = ldg(nq + at[u] / NLM);
    }
    if (b == 0) between();
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = ldg(zrows + (size_t)n[u] * NLM + (at[u] - (at[u] / NLM) * NLM));
#pragma unroll
    for (int u = 0; u < U; ++u) Z[at[u]] = (L > 64 && at[u] == one) ? 1.0 : v[u];
  }
  if (total == 0) between();  // no neighbours (single-site terms only)
  if (lane == 0) Z[(ldg(T.nb_ptr + atom + 1) - ldg(T.nb_ptr + atom)) * NLM] = 1.0;
  return atom;
}

// dz = proposed row - current row of site s, for the walk.  load() issues the
// reads before the neighbour staging (indices clamped like the staging's),
// store() writes the differences after it.
template <int L, int NLM>
struct RowDelta {
  static constexpr int U = (NLM + L - 1) / L;
  double p[U], c[U];
  __device__ __forceinline__ void load(const double* __restrict__ pz, const double* __restrict__ zs, int lane) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = lane + u * L < NLM ? lane + u * L : NLM - 1;
      p[u] = ldg(pz + k);
      c[u] = ldg(zs + k);
    }
  }
  __device__ __forceinline__ void store(double* dz, int lane) const {
#pragma unroll
    for (int u = 0; u < U; ++u) dz[lane + u * L < NLM ? lane + u * L : NLM - 1] = p[u] - c[u];
  }
};

// Orders one wavefront's LDS writes before its lanes read what other lanes
// wrote (the compiler would otherwise move accesses of different addresses
// across each other).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Fold of a group's lane partials into its first lane: the shuffle-down tree
// a += a[lane + off] for off = L/2, ..., 1.  The steps of 8 and fewer stay
// inside a 16-lane row and move data with DPP row shifts (row_shl: lane i
// reads lane i + off) instead of LDS permutes; the lanes that feed the first
// lane add the same values in the same order as the shuffle tree.
template <int CTRL>
__device__ __forceinline__ double dpp_double(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

template <int L>
__device__ __forceinline__ double fold_lanes(double a) {
#pragma unroll
  for (int off = L / 2; off >= 16; off >>= 1) a = a + __shfl_down(a, off, L);
  a = a + dpp_double<0x108>(a);
  a = a + dpp_double<0x104>(a);
  a = a + dpp_double<0x102>(a);
  a = a + dpp_double<0x101>(a);
  return a;
}

// fold_lanes<L> of the lane partials a row of LDS slots holds (slot lane of
// `row`): the shuffle steps of 32 and 16 read the partials they add straight
// from the slots, ((s[i] + s[i + 32]) + (s[i + 16] + s[i + 48])) at 64 lanes
// (s[i] + s[i + 16] at 32), the same additions in the same order for the
// lanes that feed the first; then the same DPP steps.
template <int L>
__device__ __forceinline__ double fold_slots(const double* row, int lane) {
  double a;
  if constexpr (L == 64)
    a = (row[lane] + row[(lane + 32) & 63]) + (row[(lane + 16) & 63] + row[(lane + 48) & 63]);
  else if constexpr (L == 32)
    a = row[lane] + row[(lane + 16) & 31];
  else
    a = row[lane];
  a = a + dpp_double<0x108>(a);
  a = a + dpp_double<0x104>(a);
  a = a + dpp_double<0x102>(a);
  a = a + dpp_double<0x101>(a);
  return a;
}

// ---------------------------------------------------------------------------
// the colour walk: `L` lanes per site, 256/L sites per block.
// Lane `lane` visits entries i = lane, lane + L, lane + 2L, ... in ascending
// order (the per-lane accumulation order every path and the oracle share); the
// entries are loaded WALK_UNROLL at a time ahead of their use.  (Holding the
// first entries in registers from the start of the site's update shortens its
// chain of dependent loads but costs occupancy: ~160 VGPRs, 3 waves per SIMD,
// measured 1.5x slower per sweep at 32^3.)
// ---------------------------------------------------------------------------
constexpr int WALK_UNROLL = 8;
// Every sublattice walk is padded to a multiple of WALK_PAD entries, a whole
// number of batches at every lane count, with no-op entries (weight 0, every
// factor on the 1.0 slot).
constexpr int WALK_PAD = WALK_UNROLL * 64;

// The padded walks hold LDS byte offsets from the group's staged block Z: the
// three factor slots (slot * 8) and the target row of dz, which every walk
// kernel places right after the neighbour rows and the 1.0 slot
// ((nb_max * nlm + 1 + tgt) * 8).  One SDWA add forms each LDS address from
// the block's base and a 16-bit half of the entry.
struct alignas(16) WalkWords {
  double w;
  unsigned f01, f2t;  // (f0 | f1 << 16), (f2 | tgt << 16), as byte offsets
};
typedef __attribute__((address_space(3))) const double lds_cdouble;
__device__ __forceinline__ unsigned lds_add_lo(unsigned base, unsigned packed) {
  unsigned r;
  asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0"
      : "=v"(r) : "v"(base), "v"(packed));
  return r;
}
__device__ __forceinline__ unsigned lds_add_hi(unsigned base, unsigned packed) {
  unsigned r;
  asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1"
      : "=v"(r) : "v"(base), "v"(packed));
  return r;
}
__device__ __forceinline__ double lds_at(unsigned addr) { return *(lds_cdouble*)(size_t)addr; }
typedef double __attribute__((ext_vector_type(2))) walk_f64x2;
// One walk entry, as a 16-byte global load.
__device__ __forceinline__ WalkWords ld_walk(const WalkWords* p) {
  const walk_f64x2 v = ldg(reinterpret_cast<const walk_f64x2*>(p));
  WalkWords e;
  e.w = v.x;
  e.f01 = (unsigned)__double2loint(v.y);
  e.f2t = (unsigned)__double2hiint(v.y);
  return e;
}

// The walk over a padded sublattice walk `wl` of Wp entries (Wp a multiple of
// WALK_PAD = WALK_UNROLL * 64; at 128 and 256 lanes up to L / 64 - 1 such
// pads may remain after the whole batches).
// Each lane's sum is the contract's: fma(w * p, dz[tgt], a)
// over its entries in ascending order, skipping the terms whose product p is
// zero.  The skip needs no test here: a lane's sum starts at +0 and never
// becomes -0 (a sum of values of opposite sign that cancels exactly rounds to
// +0; only an fma whose exact product is a negative number below half the
// smallest subnormal could round to -0, which unit spins' rows never reach),
// so adding the term of a zero product, (+-0) * dz, leaves the sum's bits
// unchanged for finite rows.  The padding entries add 0 * 1 * dz[0] likewise.
// Without the bounds test, the select and the index clamp, a batch is its
// loads, the LDS gathers and four FP64 operations per entry.
// Batches wholly inside the walk's leading one-factor (pair) entries, then
// its two-factor entries, skip the gathers of the 1.0 slot and the products
// with it: x * 1.0 is x exactly, so p and every sum keep their bits.
template <int L, int NLM>
__device__ __forceinline__ double walk_delta(const WalkEntry* __restrict__ wl, int Wp, int2 pre, int lane,
                                             const double* Z) {
  constexpr int B = WALK_UNROLL * L;
  const WalkWords* __restrict__ ww = reinterpret_cast<const WalkWords*>(wl) + lane;
  const unsigned zb = (unsigned)(size_t)(lds_cdouble*)Z;
  const int n1 = pre.x - pre.x % B, n2 = pre.y - pre.y % B;
  double a = 0.0;
  int base = 0;
  for (; base < n1; base += B) {
    WalkWords en[WALK_UNROLL];
    // every load of the batch is issued before the first use
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u) en[u] = ld_walk(ww + base + u * L);
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u) {
      const double p = lds_at(lds_add_lo(zb, en[u].f01));
      a = __builtin_fma(en[u].w * p, lds_at(lds_add_hi(zb, en[u].f2t)), a);
    }
  }
  for (; base < n2; base += B) {
    WalkWords en[WALK_UNROLL];
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u) en[u] = ld_walk(ww + base + u * L);
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u) {
      const double p = lds_at(lds_add_lo(zb, en[u].f01)) * lds_at(lds_add_hi(zb, en[u].f01));
      a = __builtin_fma(en[u].w * p, lds_at(lds_add_hi(zb, en[u].f2t)), a);
    }
  }
  for (; base + B <= Wp; base += B) {
    WalkWords en[WALK_UNROLL];
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u) en[u] = ld_walk(ww + base + u * L);
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u) {
      const double p = (lds_at(lds_add_lo(zb, en[u].f01)) * lds_at(lds_add_hi(zb, en[u].f01))) *
                       lds_at(lds_add_lo(zb, en[u].f2t));
      a = __builtin_fma(en[u].w * p, lds_at(lds_add_hi(zb, en[u].f2t)), a);
    }
  }
  if constexpr (L > 64) {
    // walks are padded to WALK_PAD = WALK_UNROLL * 64 entries: at L lanes up
    // to L / 64 - 1 pads of WALK_PAD / L entries per lane remain, taken in
    // ascending order
    constexpr int UT = WALK_UNROLL * 64 / L;
    static_assert(UT >= 1 && UT * L == WALK_UNROLL * 64, "a pad is whole entries per lane");
    for (; base < Wp; base += UT * L) {
      WalkWords en[UT];
#pragma unroll
      for (int u = 0; u < UT; ++u) en[u] = ld_walk(ww + base + u * L);
#pragma unroll
      for (int u = 0; u < UT; ++u) {
        const double p = (lds_at(lds_add_lo(zb, en[u].f01)) * lds_at(lds_add_hi(zb, en[u].f01))) *
                         lds_at(lds_add_lo(zb, en[u].f2t));
        a = __builtin_fma(en[u].w * p, lds_at(lds_add_hi(zb, en[u].f2t)), a);
      }
    }
  }
  return a;
}

// LDS doubles a walk group needs beyond its staged block and dz row: above
// 64 lanes, the partials of the group's second and later wavefronts for the
// fold's cross-wavefront steps
template <int L>
__host__ __device__ constexpr int walk_extra() { return L > 64 ? L - 64 : 0; }

// One group of L lanes updates one site of the colour (colour-order index q).
// Every lane of the group's wavefront calls it (wavefront-level LDS ordering).
// TEST: ΔE only, for caller-chosen sites and rows (parity hook).
template <int L, int LMAX, bool TEST>
__device__ __forceinline__ void metro_group(const DevTables& T, double* __restrict__ config,
                                            double* __restrict__ zrows,
                                            const double* __restrict__ prop,
                                            double* __restrict__ dE, int* __restrict__ acc,
                                            double beta, int q, bool valid,
                                            const int* __restrict__ test_sites, double* lds,
                                            int stride, double* __restrict__ dE_raw = nullptr) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1), PS = NLM + 4;
  const int g = threadIdx.x / L, lane = threadIdx.x % L;
  if constexpr (L >= 64) {
    // one site per wavefront (two or four at 128 or 256 lanes): its index, site and
    // sublattice tables are wave-uniform (scalar registers)
    q = __builtin_amdgcn_readfirstlane(q);
    valid = __builtin_amdgcn_readfirstlane((int)valid) != 0;
  }
  double* Z = lds + g * stride;
  double* dz = Z + T.nb_max * NLM + 1;
  int s = 0, atom = 0, W = 0;
  const WalkEntry* wl = nullptr;
  int2 pre = make_int2(0, 0);
  double u_acc = 0.0;
  if (valid) {
    s = TEST ? test_sites[q] : ldg(T.color_sites + q);
    if constexpr (L >= 64) s = __builtin_amdgcn_readfirstlane(s);
    // independent of the staging: issued with it rather than after the barrier
    const int a0 = s % T.na;
    const int w0 = ldg(T.walkp_ptr + a0);
    W = ldg(T.walkp_ptr + a0 + 1) - w0;
    wl = T.walkp + w0;
    pre = ldg_int2(T.walkp_pre + a0);
    RowDelta<L, NLM> rd;
    if constexpr (TEST) {
      rd.load(prop + (size_t)q * PS + 4, zrows + (size_t)s * NLM, lane);
      stage_neighbours<L, NLM>(T, zrows, s, lane, Z, &atom);
    } else {
      u_acc = ldg(prop + (size_t)q * PS + 3);
      atom = stage_neighbours_q<L, NLM>(T, zrows, q, s, lane, Z, [&] {
        rd.load(prop + (size_t)q * PS + 4, zrows + (size_t)s * NLM, lane);
      });
    }
    rd.store(dz, lane);
  }
  // a group's staged block is its own (L <= 64: one wavefront), so the walk
  // waits on its own wavefront's writes only, not on the block's other sites;
  // above 64 lanes the group's wavefronts meet at the block's barrier
  if constexpr (L <= 64)
    wave_lds_sync();
  else
    __syncthreads();
  double a = 0.0;
  if (valid) a = walk_delta<L, NLM>(wl, W, pre, lane, Z);
  if constexpr (L > 64) {
    // the fold's steps of +128 and +64 cross the group's wavefronts through
    // the L - 64 doubles after its dz row (xs[i] holds lane 64 + i's partial);
    // then the first wavefront folds.  At 256 lanes lane l < 64 forms
    // (a_l + a_l+128) + (a_l+64 + a_l+192), the tree's order.
    static_assert(L == 128 || L == 256, "two or four wavefronts per site");
    double* xs = dz + NLM;
    if (lane >= 64) xs[lane - 64] = a;
    __syncthreads();
    const int l0 = lane & 63;
    if constexpr (L == 128)
      a = a + xs[l0];
    else
      a = (a + xs[l0 + 64]) + (xs[l0] + xs[l0 + 128]);
    a = fold_lanes<64>(a);
  } else {
    a = fold_lanes<L>(a);
  }
  if (!valid) return;
  if (TEST) {
    if (lane == 0) dE[q] = a;
    return;
  }
  int accept = 0;
  const double* pq = prop + (size_t)q * PS;
  if (lane == 0) accept = (a <= 0.0 || u_acc < sce::det_exp(-beta * a)) ? 1 : 0;
  accept = __shfl(accept, 0, L > 64 ? 64 : L);  // above 64 lanes only the first wavefront writes
  if (accept) {
    for (int k = lane; k < NLM; k += L) zrows[(size_t)s * NLM + k] = pq[4 + k];
    if (lane == 0) {
      config[3 * (size_t)s] = pq[0];
      config[3 * (size_t)s + 1] = pq[1];
      config[3 * (size_t)s + 2] = pq[2];
    }
  }
  if (lane == 0) {
    dE[q] = accept ? a : 0.0;
    acc[q] = accept;
    if (dE_raw) dE_raw[q] = a;
  }
}

template <int L, int LMAX, bool TEST>
__global__ __launch_bounds__(TPB) void k_metro(DevTables T, double* __restrict__ config,
                                               double* __restrict__ zrows,
                                               const double* __restrict__ prop,
                                               double* __restrict__ dE, int* __restrict__ acc,
                                               const ChainDev* __restrict__ st, int q_lo,
                                               int q_n, const int* __restrict__ test_sites,
                                               double* __restrict__ dE_raw = nullptr) {
  constexpr int G = TPB / L;
  extern __shared__ double lds[];
  const int qi = xcd_swizzle(blockIdx.x, gridDim.x) * G + (int)threadIdx.x / L;
  const double beta = TEST ? 0.0 : st->beta;
  metro_group<L, LMAX, TEST>(T, config, zrows, prop, dE, acc, beta, q_lo + qi, qi < q_n,
                             test_sites, lds,
                             T.nb_max * (LMAX + 1) * (LMAX + 1) + 1 + (LMAX + 1) * (LMAX + 1) + walk_extra<L>(),
                             dE_raw);
}

// ---------------------------------------------------------------------------
// overrelaxation (updates.jl:66-87): leave-one-out coefficients c by the same
// walk (lane partials per target row, shuffle-tree fold), reflection about the
// l = 1 axis h = (c[3], c[1], c[2]), exact ΔE = Σ_k c_k (znew_k - zold_k)
// TEST: writes c only.
// ---------------------------------------------------------------------------
// SLOTS: the row sums accumulate in lane-private LDS slots (row-major, one
// per lane and row after the group's staged block) with LDS adds, each entry
// added to its row's slot, over the padded walk; otherwise in registers, every
// entry's term selected into its row.  Each lane's slot of a row takes that
// lane's terms in walk order either way (the LDS applies a wavefront's
// operations in order), so both give the same sums.
template <int L, int LMAX, bool TEST, bool SLOTS>
__global__ __launch_bounds__(TPB) void k_or(DevTables T, double* __restrict__ config,
                                            double* __restrict__ zrows, double* __restrict__ dE,
                                            int* __restrict__ acc, int* __restrict__ att,
                                            const ChainDev* __restrict__ st, int or_pass,
                                            int q_lo, int q_n, const int* __restrict__ test_sites,
                                            double* __restrict__ c_out) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1), G = TPB / L;
  extern __shared__ double lds[];
  const int g = threadIdx.x / L, lane = threadIdx.x % L;
  int qi = xcd_swizzle(blockIdx.x, gridDim.x) * G + g;
  // one site per wavefront at 64 lanes: its index and site are wave-uniform
  if constexpr (L == 64) qi = __builtin_amdgcn_readfirstlane(qi);
  const int q = q_lo + qi;
  int s = 0, atom = 0;
  bool valid = qi < q_n;
  if (valid) {
    s = TEST ? test_sites[q] : T.color_sites[q];
    if constexpr (L == 64) s = __builtin_amdgcn_readfirstlane(s);
    if (!TEST && !T.has_l1[s]) {
      if (lane == 0) {
        dE[q] = 0.0;
        acc[q] = 0;
        att[q] = 0;
      }
      valid = false;
    }
  }
  const int stride = T.nb_max * NLM + 1 + NLM + (SLOTS ? L * NLM : 0);
  double* Z = lds + g * stride;
  double* slot = Z + T.nb_max * NLM + 1 + NLM;  // SLOTS: slot[k * L + lane]
  // the site's own spin and row and its keyed accept uniform do not depend on
  // the walk: read and drawn while the neighbour rows stage
  double e0[3] = {0.0, 0.0, 0.0}, zo0[NLM], u_or = 0.0;
#pragma unroll
  for (int k = 0; k < NLM; ++k) zo0[k] = 0.0;
  if (valid) {
    if constexpr (TEST) {
      stage_neighbours<L, NLM>(T, zrows, s, lane, Z, &atom);
    } else {
      atom = stage_neighbours_q<L, NLM>(T, zrows, q, s, lane, Z, [&] {
        if (lane == 0) {
          e0[0] = config[3 * (size_t)s];
          e0[1] = config[3 * (size_t)s + 1];
          e0[2] = config[3 * (size_t)s + 2];
#pragma unroll
          for (int k = 0; k < NLM; ++k) zo0[k] = zrows[(size_t)s * NLM + k];
          const sce::u32x4 b = sce::philox_block(st->seed, s + 1, (int32_t)st->sweep_index,
                                                 sce::SLOT_OR0 + (uint32_t)or_pass, 0u);
          u_or = sce::uniform01(b.x, b.y);
        }
      });
    }
  }
  double c[NLM];
#pragma unroll
  for (int k = 0; k < NLM; ++k) c[k] = 0.0;
  if constexpr (SLOTS) {
#pragma unroll
    for (int k = 0; k < NLM; ++k) slot[k * L + lane] = 0.0;
  }
  wave_lds_sync();  // per-group LDS: the own wavefront only
  if (SLOTS && valid) {
    // the padded walk (no-op padding, pair / two-factor prefixes as in
    // walk_delta); an entry's target byte offset (dz0 + k) * 8 becomes its
    // slot (k * L + lane) * 8: ((off - dz0 * 8) << log2 L) + lane * 8
    constexpr int B = WALK_UNROLL * L, SH = L == 16 ? 4 : (L == 32 ? 5 : 6);
    const int w0 = T.walkp_ptr[atom];
    const int Wp = T.walkp_ptr[atom + 1] - w0;
    const int2 pre = T.walkp_pre[atom];
    const WalkWords* __restrict__ ww = reinterpret_cast<const WalkWords*>(T.walkp + w0) + lane;
    const unsigned zb = (unsigned)(size_t)(lds_cdouble*)Z;
    const unsigned sb = (unsigned)(size_t)(lds_cdouble*)slot + (unsigned)lane * 8u -
                        ((unsigned)(T.nb_max * NLM + 1) * 8u << SH);
    typedef __attribute__((address_space(3))) double lds_double;
    auto add_to_row = [&](unsigned f2t, double v) {
      lds_double* a = (lds_double*)(size_t)(sb + ((f2t >> 16) << SH));
      __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    };
    const int n1 = pre.x - pre.x % B, n2 = pre.y - pre.y % B;
    int base = 0;
    for (; base < n1; base += B) {
      WalkWords en[WALK_UNROLL];
#pragma unroll
      for (int u = 0; u < WALK_UNROLL; ++u) en[u] = ld_walk(ww + base + u * L);
#pragma unroll
      for (int u = 0; u < WALK_UNROLL; ++u) add_to_row(en[u].f2t, en[u].w * lds_at(lds_add_lo(zb, en[u].f01)));
    }
    for (; base < n2; base += B) {
      WalkWords en[WALK_UNROLL];
#pragma unroll
      for (int u = 0; u < WALK_UNROLL; ++u) en[u] = ld_walk(ww + base + u * L);
#pragma unroll
      for (int u = 0; u < WALK_UNROLL; ++u)
        add_to_row(en[u].f2t, en[u].w * (lds_at(lds_add_lo(zb, en[u].f01)) * lds_at(lds_add_hi(zb, en[u].f01))));
    }
    for (; base < Wp; base += B) {
      WalkWords en[WALK_UNROLL];
#pragma unroll
      for (int u = 0; u < WALK_UNROLL; ++u) en[u] = ld_walk(ww + base + u * L);
#pragma unroll
      for (int u = 0; u < WALK_UNROLL; ++u)
        add_to_row(en[u].f2t, en[u].w * ((lds_at(lds_add_lo(zb, en[u].f01)) * lds_at(lds_add_hi(zb, en[u].f01))) *
                                         lds_at(lds_add_lo(zb, en[u].f2t))));
    }
  }
  if constexpr (SLOTS) {
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < NLM; ++k) c[k] = fold_slots<L>(slot + k * L, lane);
  }
  if (!SLOTS && valid) {
    const int w0 = T.walk_ptr[atom];
    const int W = T.walk_ptr[atom + 1] - w0;
    const WalkEntry* __restrict__ wl = T.walk + w0;
    for (int base = 0; base < W; base += WALK_UNROLL * L) {
      WalkEntry en[WALK_UNROLL];
#pragma unroll
      for (int u = 0; u < WALK_UNROLL; ++u) {
        const int i = base + u * L + lane;
        en[u] = wl[i < W ? i : W - 1];  // clamped: the batch's loads issue together
      }
#pragma unroll
      for (int u = 0; u < WALK_UNROLL; ++u) {
        const int i = base + u * L + lane;
        const double p = (Z[en[u].f0] * Z[en[u].f1]) * Z[en[u].f2];
        const bool use = i < W && p != 0.0;
        const double v = en[u].w * p;
#pragma unroll
        for (int k = 0; k < NLM; ++k) c[k] = (use && k == en[u].tgt) ? c[k] + v : c[k];
      }
    }
  }
  // per row, the shuffle-down tree into the first lane (DPP for its last steps)
  if constexpr (!SLOTS) {
#pragma unroll
    for (int k = 0; k < NLM; ++k) c[k] = fold_lanes<L>(c[k]);
  }
  if (!valid || lane != 0) return;
  if (TEST) {
#pragma unroll
    for (int k = 0; k < NLM; ++k) c_out[(size_t)q * NLM + k] = c[k];
    return;
  }
  int at = 0, ac = 0;
  double d_e = 0.0;
  if constexpr (LMAX >= 1) {
    const double h0 = c[3], h1 = c[1], h2 = c[2];
    const double hn = __builtin_sqrt((h0 * h0 + h1 * h1) + h2 * h2);
    if (!(hn < 1e-12)) {
      at = 1;
      const double ax[3] = {h0 / hn, h1 / hn, h2 / hn};
      const double* e = e0;
      const double dd = (e[0] * ax[0] + e[1] * ax[1]) + e[2] * ax[2];
      const double tt = 2.0 * dd;
      const double e2[3] = {tt * ax[0] - e[0], tt * ax[1] - e[1], tt * ax[2] - e[2]};
      double zn[NLM];
      sce::zlm_row<LMAX>(e2[0], e2[1], e2[2], zn);
      double dEv = 0.0;
#pragma unroll
      for (int k = 0; k < NLM; ++k) {
        if (c[k] == 0.0) continue;
        dEv += c[k] * (zn[k] - zo0[k]);
      }
      if (dEv <= 0.0 || u_or < sce::det_exp(-st->beta * dEv)) {
        ac = 1;
        d_e = dEv;
        double* ew = config + 3 * (size_t)s;
        ew[0] = e2[0];
        ew[1] = e2[1];
        ew[2] = e2[2];
        double* zw = zrows + (size_t)s * NLM;
#pragma unroll
        for (int k = 0; k < NLM; ++k) zw[k] = zn[k];
      }
    }
  }
  dE[q] = d_e;
  acc[q] = ac;
  att[q] = at;
}

// ---------------------------------------------------------------------------
// deterministic tree reduction: chunks of 512 (zero padded) folded by offsets
// 256..1; a multi-chunk input recurses over the chunk sums.
// ---------------------------------------------------------------------------
enum { MODE_METRO = 0, MODE_OR = 1, MODE_RENORM = 2, MODE_TOTAL = 3, MODE_MEASURE = 4, MODE_TEST = 5,
       MODE_DA = 6 };

__device__ double block_tree(double x0, double x1, double* sh) {
  // 256 threads, v[t] = x0 + x1 is the offset-256 step
  const int t = threadIdx.x;
  double v = x0 + x1;
  sh[t] = v;
  __syncthreads();
  if (t < 128) sh[t] = v = v + sh[t + 128];
  __syncthreads();
  if (t < 64) v = v + sh[t + 64];
  if (t < 64) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_down(v, off, 64);
  }
  __syncthreads();
  return v;  // valid in thread 0
}

// ---------------------------------------------------------------------------
// Streaming error analysis on the device: the LogBinner (src/binning.jl:25-79)
// and BinStore (:138-188) of the five standard series of a measurement row
// (E, m): E, E^2, |m|, m^2, m^4 with m^2 = (m_x^2 + m_y^2) + m_z^2, |m| its
// square root, m^4 = m^2 m^2.  The measurement's final reduction thread pushes
// each row, so a run's statistics need no download of the rows.
// ---------------------------------------------------------------------------
constexpr int BIN_COMP = 5, BIN_LEVELS = 62, BIN_MAXSTORE = 1024;

struct DevBinner {
  long long n;
  long long count[BIN_LEVELS];
  int pending_full[BIN_LEVELS];
  double sums[BIN_LEVELS][BIN_COMP], sums2[BIN_LEVELS][BIN_COMP], pending[BIN_LEVELS][BIN_COMP];
  int bin_size, nbins, nfull, nacc;
  double acc[BIN_COMP];
  double means[BIN_MAXSTORE][BIN_COMP];
};

__device__ void binner_push(DevBinner* b, double E, double mx, double my, double mz) {
  const double m2 = (mx * mx + my * my) + mz * mz;
  const double x[BIN_COMP] = {E, E * E, __builtin_sqrt(m2), m2, m2 * m2};
  double cur[BIN_COMP];
#pragma unroll
  for (int j = 0; j < BIN_COMP; ++j) cur[j] = x[j];
  b->n += 1;
  for (int k = 0; k < BIN_LEVELS; ++k) {
    b->count[k] += 1;
#pragma unroll
    for (int j = 0; j < BIN_COMP; ++j) {
      b->sums[k][j] = b->sums[k][j] + cur[j];
      b->sums2[k][j] = b->sums2[k][j] + cur[j] * cur[j];
    }
    if (b->pending_full[k]) {
#pragma unroll
      for (int j = 0; j < BIN_COMP; ++j) {
        const double pj = 0.5 * (b->pending[k][j] + cur[j]);
        b->pending[k][j] = pj;
        cur[j] = pj;
      }
      b->pending_full[k] = 0;
    } else {
#pragma unroll
      for (int j = 0; j < BIN_COMP; ++j) b->pending[k][j] = cur[j];
      b->pending_full[k] = 1;
      break;
    }
  }
  if (b->bin_size > 0 && b->nfull < b->nbins) {
#pragma unroll
    for (int j = 0; j < BIN_COMP; ++j) b->acc[j] = b->acc[j] + x[j];
    if (++b->nacc == b->bin_size) {
#pragma unroll
      for (int j = 0; j < BIN_COMP; ++j) {
        b->means[b->nfull][j] = b->acc[j] / (double)b->bin_size;
        b->acc[j] = 0.0;
      }
      b->nfull += 1;
      b->nacc = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// fence-free two-kernel reduction: k_chunks writes the 512-element chunk trees
// (and integer counts, and optionally the three mean-spin components); one
// block of k_final folds them (the kernel boundary makes the partials visible
// across XCDs, so no device-scope fence is needed) and updates the chain.
// ---------------------------------------------------------------------------
// One block's share of the first reduction level: chunk b of x (and of the
// three spin components when meas).
template <int MODE>
__device__ void chunk_body(long long b, const double* __restrict__ x, long long n,
                           const int* __restrict__ iacc, const int* __restrict__ iatt,
                           double* __restrict__ part, long long* __restrict__ ipart,
                           const double* __restrict__ config, const uint8_t* __restrict__ active,
                           long long n_sites, int meas) {
  __shared__ double sh[256];
  __shared__ long long csum[8];
  const int t = threadIdx.x;
  const long long nbx = (n + TREE_B - 1) / TREE_B > 0 ? (n + TREE_B - 1) / TREE_B : 1;
  const long long nchm = (n_sites + TREE_B - 1) / TREE_B;
  if (b < nbx) {
    const long long i0 = b * TREE_B + t, i1 = i0 + 256;
    const double v = block_tree(i0 < n ? x[i0] : 0.0, i1 < n ? x[i1] : 0.0, sh);
    if (t == 0) part[b] = v;
    if (MODE == MODE_METRO || MODE == MODE_OR || MODE == MODE_DA) {
      long long ca = (i0 < n ? iacc[i0] : 0) + (i1 < n ? iacc[i1] : 0);
      long long ct = 0;
      if (MODE == MODE_OR || MODE == MODE_DA) ct = (i0 < n ? iatt[i0] : 0) + (i1 < n ? iatt[i1] : 0);
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        ca += __shfl_down(ca, off, 64);
        ct += __shfl_down(ct, off, 64);
      }
      if ((t & 63) == 0) {
        csum[t >> 6] = ca;
        csum[4 + (t >> 6)] = ct;
      }
      __syncthreads();
      if (t == 0) {
        ipart[2 * b] = csum[0] + csum[1] + csum[2] + csum[3];
        ipart[2 * b + 1] = csum[4] + csum[5] + csum[6] + csum[7];
      }
    }
  }
  if (meas && b < nchm) {
    const long long i0 = b * TREE_B + t, i1 = i0 + 256;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double x0 = (i0 < n_sites && active[i0]) ? config[3 * i0 + c] : 0.0;
      const double x1 = (i1 < n_sites && active[i1]) ? config[3 * i1 + c] : 0.0;
      const double v = block_tree(x0, x1, sh);
      if (t == 0) part[nbx + c * nchm + b] = v;
    }
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_chunks(const double* __restrict__ x, long long n,
                                                const int* __restrict__ iacc,
                                                const int* __restrict__ iatt,
                                                double* __restrict__ part,
                                                long long* __restrict__ ipart,
                                                const double* __restrict__ config,
                                                const uint8_t* __restrict__ active,
                                                long long n_sites, int meas) {
  chunk_body<MODE>(blockIdx.x, x, n, iacc, iatt, part, ipart, config, active, n_sites, meas);
}

// The first reduction level of a Metropolis sweep fused with proposals of the
// next sweep: blocks below nchunk reduce, the others propose the range r (the
// sweep's walks are complete, so every spin is final for the sweep).
template <int LMAX, int MODE = MODE_METRO>
__global__ __launch_bounds__(256) void k_chunks_propose(const double* __restrict__ x, long long n,
                                                        const int* __restrict__ iacc, const int* __restrict__ iatt,
                                                        double* __restrict__ part,
                                                        long long* __restrict__ ipart,
                                                        const double* __restrict__ config,
                                                        const uint8_t* __restrict__ active, long long n_sites,
                                                        int meas, long long nchunk, DevTables T,
                                                        double* __restrict__ prop, const ChainDev* __restrict__ st,
                                                        PropRange r) {
  if ((long long)blockIdx.x < nchunk) {
    chunk_body<MODE>(blockIdx.x, x, n, iacc, iatt, part, ipart, config, active, n_sites, meas);
    return;
  }
  const int t = (int)(((long long)blockIdx.x - nchunk) * blockDim.x + threadIdx.x);
  propose_body<LMAX>(T, config, prop, st, r.lo, r.n, t, r.off);
}

__device__ double final_tree(const double* buf, long long m, double* scratch, double* sh);

template <int MODE>
__device__ void final_body(double* part, long long n, const long long* ipart, long long n_sites,
                           int meas, ChainDev* st, double* __restrict__ out, long long n_att,
                           double n_div, double* scratch, DevBinner* bin = nullptr) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  const long long nbx = (n + TREE_B - 1) / TREE_B > 0 ? (n + TREE_B - 1) / TREE_B : 1;
  const long long nchm = (n_sites + TREE_B - 1) / TREE_B;
  double sum;
  double m3[3] = {0.0, 0.0, 0.0};
  long long nacc = 0, natt = 0;
  if (MODE == MODE_METRO || MODE == MODE_OR || MODE == MODE_DA) {
    // integer counts: order-free sum over the chunk partials
    long long ca = 0, ct = 0;
    for (long long b = t; b < nbx; b += 256) {
      ca += ipart[2 * b];
      ct += ipart[2 * b + 1];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      ca += __shfl_down(ca, off, 64);
      ct += __shfl_down(ct, off, 64);
    }
    __shared__ long long cs[8];
    if ((t & 63) == 0) {
      cs[t >> 6] = ca;
      cs[4 + (t >> 6)] = ct;
    }
    __syncthreads();
    nacc = cs[0] + cs[1] + cs[2] + cs[3];
    natt = cs[4] + cs[5] + cs[6] + cs[7];
  }
  if (nbx > 1 && nbx <= TREE_B && (!meas || nchm <= TREE_B)) {
    // single-level case: every component's partials loaded in one pass
    double a0[4], a1[4];
    const long long len[4] = {nbx, nchm, nchm, nchm};
    const double* src[4] = {part, part + nbx, part + nbx + nchm, part + nbx + 2 * nchm};
    const int nc = meas ? 4 : 1;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      a0[c] = (c < nc && t < len[c]) ? src[c][t] : 0.0;
      a1[c] = (c < nc && t + 256 < len[c]) ? src[c][t + 256] : 0.0;
    }
    if (nchm == 1) {  // a single chunk is already the final value
      a0[1] = t == 0 ? src[1][0] : 0.0;
      a0[2] = t == 0 ? src[2][0] : 0.0;
      a0[3] = t == 0 ? src[3][0] : 0.0;
    }
    sum = block_tree(a0[0], a1[0], sh);
    if (meas) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double v = (nchm == 1) ? a0[c + 1] : block_tree(a0[c + 1], a1[c + 1], sh);
        m3[c] = v;
      }
    }
  } else {
    sum = final_tree(part, nbx, scratch, sh);
    if (meas) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        __syncthreads();
        m3[c] = final_tree(part + nbx + c * nchm, nchm, scratch, sh);
      }
    }
  }
  if (t != 0) return;
  if (MODE == MODE_METRO || MODE == MODE_DA) {
    st->energy = st->energy + sum;
    st->acc_metro += nacc;
    st->att_metro += n_att;
    st->nacc_last = nacc;
    st->sweep_index += 1;
    if (MODE == MODE_DA) st->acc_stage1 += natt;
  } else if (MODE == MODE_OR) {
    st->energy = st->energy + sum;
    st->acc_or += nacc;
    st->att_or += natt;
    st->nacc_last = nacc;
  } else if (MODE == MODE_RENORM) {
    const double drift = fabs(st->energy - sum);
    st->last_drift = drift;
    if (drift > st->max_drift) st->max_drift = drift;
    st->energy = sum;
  } else if (MODE == MODE_TOTAL) {
    st->e_full = sum;
  } else if (MODE == MODE_TEST) {
    out[0] = sum;
  }
  if (meas) {
    const int row = st->n_obs;
    const double E = st->energy, mx = m3[0] / n_div, my = m3[1] / n_div, mz = m3[2] / n_div;
    out[4 * (size_t)row] = E;
    out[4 * (size_t)row + 1] = mx;
    out[4 * (size_t)row + 2] = my;
    out[4 * (size_t)row + 3] = mz;
    st->n_obs = row + 1;
    if (bin) binner_push(bin, E, mx, my, mz);
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_final(double* part, long long n, const long long* ipart,
                                               long long n_sites, int meas, ChainDev* st,
                                               double* __restrict__ out, long long n_att,
                                               double n_div, double* scratch, DevBinner* bin) {
  final_body<MODE>(part, n, ipart, n_sites, meas, st, out, n_att, n_div, scratch, bin);
}

// The previous sweep's final reduction (k_final<MODE_METRO> of that sweep),
// deferred into block 0 of the next sweep's walk of its smallest colour class:
// its operands (the chunk partials and counts) are complete at that launch's
// start and are rewritten only by the sweep's own reduction after it, and no
// walk block reads what it writes (the walk reads only st->beta).  Same
// arithmetic as the standalone launch, one kernel fewer per sweep.  The final
// needs more registers than the walk (fewer waves per SIMD), which the
// smallest class does not feel.
struct FinalArgs {
  double* part;
  const long long* ipart;
  double* out;
  double* scratch;
  DevBinner* bin;
  long long n, n_sites, n_att;
  double n_div;
  int meas;
};

template <int L, int LMAX>
__global__ __launch_bounds__(TPB) void k_metro_fin(DevTables T, double* __restrict__ config,
                                                   double* __restrict__ zrows, const double* __restrict__ prop,
                                                   double* __restrict__ dE, int* __restrict__ acc, ChainDev* st,
                                                   int q_lo, int q_n, FinalArgs f) {
  if (blockIdx.x == 0) {
    final_body<MODE_METRO>(f.part, f.n, f.ipart, f.n_sites, f.meas, st, f.out, f.n_att, f.n_div, f.scratch, f.bin);
    return;
  }
  constexpr int G = TPB / L, NLM = (LMAX + 1) * (LMAX + 1);
  extern __shared__ double lds[];
  const int qi = xcd_swizzle(blockIdx.x - 1, gridDim.x - 1) * G + (int)threadIdx.x / L;
  metro_group<L, LMAX, false>(T, config, zrows, prop, dE, acc, st->beta, q_lo + qi, qi < q_n, nullptr, lds,
                              T.nb_max * NLM + 1 + NLM + walk_extra<L>());
}

// A colour walk whose extra blocks (after the nwalk walk blocks) make
// proposals of other colour classes: r0 then r1, each in whole blocks.  The
// proposals read the spins of classes whose walks are complete and write
// their own classes' proposal records; the walk reads only its own class's.
template <int L, int LMAX>
__global__ __launch_bounds__(TPB) void k_metro_prop(DevTables T, double* __restrict__ config,
                                                    double* __restrict__ zrows, double* __restrict__ prop,
                                                    double* __restrict__ dE, int* __restrict__ acc,
                                                    const ChainDev* __restrict__ st, int q_lo, int q_n, int nwalk,
                                                    PropRange r0, PropRange r1) {
  if ((int)blockIdx.x >= nwalk) {
    const int b = (int)blockIdx.x - nwalk, nb0 = prop_blocks(r0.n);
    if (b < nb0)
      propose_body<LMAX>(T, config, prop, st, r0.lo, r0.n, b * TPB + (int)threadIdx.x, r0.off);
    else
      propose_body<LMAX>(T, config, prop, st, r1.lo, r1.n, (b - nb0) * TPB + (int)threadIdx.x, r1.off);
    return;
  }
  constexpr int G = TPB / L, NLM = (LMAX + 1) * (LMAX + 1);
  extern __shared__ double lds[];
  const int qi = xcd_swizzle(blockIdx.x, nwalk) * G + (int)threadIdx.x / L;
  metro_group<L, LMAX, false>(T, config, zrows, prop, dE, acc, st->beta, q_lo + qi, qi < q_n, nullptr, lds,
                              T.nb_max * NLM + 1 + NLM + walk_extra<L>());
}

// Block-level recursive tree over m partial sums (the levels above the first
// of ora-style tree: m <= 512 -> one chunk tree, else chunk sums recursively).
// Valid in thread 0.  `scratch` receives intermediate levels.
__device__ double final_tree(const double* buf, long long m, double* scratch, double* sh) {
  volatile const double* vb = buf;
  if (m == 1) return vb[0];
  while (true) {
    const int t = threadIdx.x;
    if (m <= TREE_B) {
      const double a0 = t < m ? vb[t] : 0.0;
      const double a1 = t + 256 < m ? vb[t + 256] : 0.0;
      return block_tree(a0, a1, sh);
    }
    const long long nb = (m + TREE_B - 1) / TREE_B;
    volatile double* vs = scratch;
    for (long long b = 0; b < nb; ++b) {
      const long long j0 = b * TREE_B + t, j1 = j0 + 256;
      const double a0 = j0 < m ? vb[j0] : 0.0;
      const double a1 = j1 < m ? vb[j1] : 0.0;
      const double v = block_tree(a0, a1, sh);
      if (t == 0) vs[b] = v;
      __syncthreads();
    }
    vb = scratch;
    scratch += nb;
    m = nb;
  }
}

template <int LMAX>
__device__ __forceinline__ void propose_one(const DevTables& T, const double* __restrict__ config,
                                            double* __restrict__ prop, int q, uint64_t seed,
                                            int32_t sweep, double step) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1), PS = NLM + 4;
  const int s = T.color_sites[q];
  const double e[3] = {config[3 * (size_t)s], config[3 * (size_t)s + 1], config[3 * (size_t)s + 2]};
  double e2[3], u, z[NLM];
  sce::keyed_proposal(seed, s + 1, sweep, e, step, e2, &u);
  sce::zlm_row<LMAX>(e2[0], e2[1], e2[2], z);
  double* o = prop + (size_t)q * PS;
  o[0] = e2[0];
  o[1] = e2[1];
  o[2] = e2[2];
  o[3] = u;
#pragma unroll
  for (int k = 0; k < NLM; ++k) o[4 + k] = z[k];
}

// ---------------------------------------------------------------------------
// Dependency-driven persistent sweep: one cooperative launch runs `nsweeps`
// Metropolis sweeps with no grid barrier.  Wave w owns the colour positions
// q = w, w + W, ... and updates them for sweep 1, 2, ... in (sweep, position)
// order.  The colour-serial schedule fixes what each update reads: a
// neighbour of lower colour after its update of the same sweep, one of higher
// colour after its update of the previous sweep (and before its update of this
// one, which waits on this site in turn).  So item (t, q) waits until every
// neighbour n has done[n] >= t (lower colour) or >= t - 1 (higher colour;
// padding slots hold the site itself), and publishes done[s] = t once its
// writes are complete: every update reads and writes exactly what it does in
// the per-colour launches, so the results are the same bits.  Every
// dependency of an item comes earlier in (sweep, position) order and every
// wave walks its items in that order, so with all waves resident the earliest
// unfinished item can always proceed: no deadlock.  Waits are bounded anyway
// (spin limit -> *err, every wave leaves).
//
// Data one wave writes and another reads (rows, spins, proposals, done flags)
// moves through agent-scope atomic loads and stores: coherent across the XCDs'
// separate L2 caches without cache-wide write-backs or invalidations.  A
// wave-scope release (wait for its stores) orders an update's writes before
// its done flag; a wave-scope acquire orders a waiting wave's reads after the
// flags it saw.  Per-sweep outputs go to segment buffers (dE and accept flags
// by colour position, the spins of measured sweeps by site), reduced after
// the launch by the same chunk and final trees as the per-colour path.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ld_ag(const double* p) {
  return __longlong_as_double(
      __hip_atomic_load(reinterpret_cast<const long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void st_ag(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<long long*>(p), __double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_ag(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_ag(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr unsigned FLOW_SPIN_LIMIT = 1u << 22;

struct FlowArgs {
  DevTables T;
  double *config, *zrows, *prop;
  double* dE_seg;    // nsweeps x n_active, colour-position order
  int* acc_seg;      // nsweeps x n_active
  double* cfg_seg;   // measured sweeps x 3 x n_sites (spins after the sweep)
  int* done;         // per site: the last sweep (1-based, absolute) it completed
  const unsigned long long* low;  // per position: bit j = neighbour slot j has a lower colour
  int* err;
  long long t0;      // completed sweeps before the launch
  int nsweeps;
  long long i0;      // sweeps of this call before the launch (the measurement schedule)
  int measure_interval;
  int defer_last;    // the last sweep's measurement follows a renormalisation
  double beta, step;
  unsigned long long seed;
  unsigned spin_limit;
};

// measurement slot of segment sweep i (1-based) or -1
__host__ __device__ inline int flow_meas_slot(long long i0, int i, int n, int mi, int defer_last) {
  if (mi <= 0 || ((i0 + i) % mi) != 0 || (defer_last && i == n)) return -1;
  return (int)((i0 + i) / mi - i0 / mi - 1);
}

// One item of a dependency-driven sweep per group of L lanes (64 / L sites of
// one colour class per wavefront, never neighbours of each other): wait until
// every neighbour has done[nb] >= t (lower colour) or >= t - 1 (otherwise),
// stage the neighbour rows, walk, decide, write the site, its sweep outputs
// (dE and the accept flag; the spin when cfg_out) and publish done[s] = t.
// nb[m]: this lane's neighbour slot lane + m L (the site itself past nb_max).
// A group with `valid` false mirrors a valid site of its wave and writes
// nothing.  `next` runs between the staging and the walk (the caller's
// prefetch of its next item).  Called by every lane of the wave; false when a
// wait exceeded its bound (the wave must leave).
template <int LMAX, int L, class Next>
__device__ __forceinline__ bool flow_update(const DevTables& T, double* config, double* zrows, const double* prop,
                                            int* done, int* err, unsigned spin_limit, int q, int s,
                                            unsigned long long low, const int (&nb)[64 / L], bool valid, int t,
                                            double beta, double* Z, double* dz, double* dE_out, int* acc_out,
                                            double* cfg_out, Next&& next) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1), PS = NLM + 4, NSL = 64 / L;
  static_assert(NLM <= L, "one lane per row element");
  const int lane = threadIdx.x % L;
  const int nbm = T.nb_max;
  const int k = lane < NLM ? lane : NLM - 1;
  // this site's proposal (made by this wave), current row and spin: the
  // wave's own earlier writes, independent of the neighbours' progress
  __builtin_amdgcn_s_waitcnt(0);
  const double* pq = prop + (size_t)q * PS;
  const double u_acc = ld_ag(pq + 3);
  const double p_new = ld_ag(pq + 4 + k), p_cur = ld_ag(zrows + (size_t)s * NLM + k);
  const double e_old = lane < 3 ? ld_ag(config + 3 * (size_t)s + lane) : 0.0;
  const double e_new = lane < 3 ? ld_ag(pq + lane) : 0.0;
  // wait for the neighbours (lane j: slots j, j + L, ...)
  {
    int need[NSL];
#pragma unroll
    for (int m = 0; m < NSL; ++m) {
      const int j = lane + m * L;
      need[m] = (j < nbm && ((low >> j) & 1ull)) ? t : t - 1;
    }
    unsigned spins = 0;
    while (true) {
      bool ok = true;
#pragma unroll
      for (int m = 0; m < NSL; ++m) ok = ok && ld_ag(done + nb[m]) >= need[m];
      if (__all(ok)) break;
      if (++spins > spin_limit) {
        if (threadIdx.x == 0) st_ag(err, 1);
        return false;
      }
      if ((spins & 63u) == 0 && ld_ag(err) != 0) return false;
      if (spins > 4) __builtin_amdgcn_s_sleep(1);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // stage the neighbour rows (slot-major, nb_max rows) and the 1.0 slot
  const int32_t* nq = T.nbq + (size_t)q * nbm;
  const int total = nbm * NLM;
  for (int b = 0; b < total; b += 4 * L) {
    int at[4];
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int x = b + u * L + lane;
      at[u] = x < total ? x : total - 1;
      const int nbx = nq[at[u] / NLM];
      v[u] = ld_ag(zrows + (size_t)nbx * NLM + (at[u] - (at[u] / NLM) * NLM));
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) Z[at[u]] = v[u];
  }
  const int atom = s % T.na;
  if (lane == 0) Z[(T.nb_ptr[atom + 1] - T.nb_ptr[atom]) * NLM] = 1.0;
  if (lane < NLM) dz[lane] = p_new - p_cur;
  next();
  __syncthreads();
  const int w0 = T.walkp_ptr[atom];
  double dsum = walk_delta<L, NLM>(T.walkp + w0, T.walkp_ptr[atom + 1] - w0, T.walkp_pre[atom], lane, Z);
  dsum = fold_lanes<L>(dsum);
  int accept = 0;
  if (lane == 0) accept = (dsum <= 0.0 || u_acc < sce::det_exp(-beta * dsum)) ? 1 : 0;
  accept = __shfl(accept, 0, L);
  if (valid) {
    if (accept) {
      if (lane < NLM) st_ag(zrows + (size_t)s * NLM + lane, p_new);
      if (lane < 3) st_ag(config + 3 * (size_t)s + lane, e_new);
    }
    const double e_now = accept ? e_new : e_old;
    if (cfg_out && lane < 3) st_ag(cfg_out + lane, e_now);
    if (lane == 0) {
      st_ag(dE_out, accept ? dsum : 0.0);
      st_ag(acc_out, accept);
    }
  }
  // publish: this update's writes complete before its flag
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0);
  if (valid && lane == 0) st_ag(done + s, t);
  __syncthreads();  // LDS reused by the next item
  return true;
}

// The proposals of sweep `sweep` for the positions q0, q0 + W, ... < n, from
// their current spins (written by this wave only): sixteen positions at a
// time, four lanes each.
template <int LMAX>
__device__ __forceinline__ void flow_propose(const DevTables& T, const double* config, double* prop, int q0, int W,
                                             unsigned long long seed, int32_t sweep, double step) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1), PS = NLM + 4;
  const int lane = threadIdx.x, n = T.n_active;
  const int g = lane / PROPOSE_LANES;
  for (int qb = q0; qb < n; qb += 16 * W) {
    const int qg = qb + g * W;
    const int qv = qg < n ? qg : qb;  // every lane takes part in the exchange
    const int sv = T.color_sites[qv];
    const double e[3] = {ld_ag(config + 3 * (size_t)sv), ld_ag(config + 3 * (size_t)sv + 1),
                         ld_ag(config + 3 * (size_t)sv + 2)};
    double e2[3], u = 0.0, z[NLM];
    if (propose_core<LMAX>(seed, sv, sweep, e, step, lane % PROPOSE_LANES, e2, &u, z) && qg < n) {
      double* o = prop + (size_t)qg * PS;
      st_ag(o, e2[0]);
      st_ag(o + 1, e2[1]);
      st_ag(o + 2, e2[2]);
      st_ag(o + 3, u);
#pragma unroll
      for (int kk = 0; kk < NLM; ++kk) st_ag(o + 4 + kk, z[kk]);
    }
  }
}

// The proposals of sweep `sweep` for the positions of units u0, u0 + W, ...
// < nu (unit u: positions ulo[u] .. ulo[u + 1] - 1, at most G of them): as
// flow_propose, sixteen positions at a time.
template <int LMAX>
__device__ __forceinline__ void flow_propose_units(const DevTables& T, const double* config, double* prop,
                                                   const int* ulo, int nu, int u0, int W, int G,
                                                   unsigned long long seed, int32_t sweep, double step) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1), PS = NLM + 4;
  const int lane = threadIdx.x;
  const int nf = ((nu - u0 + W - 1) / W) * G;  // position slots of this wave's units
  for (int fb = 0; fb < nf; fb += 16) {
    const int f = fb + lane / PROPOSE_LANES;
    const int u = u0 + (f / G) * W;
    const int qg = u < nu ? ulo[u] + f % G : -1;
    const bool ok = u < nu && f < nf && qg < ulo[u + 1];
    const int qv = ok ? qg : ulo[u0];  // every lane takes part in the exchange
    const int sv = T.color_sites[qv];
    const double e[3] = {ld_ag(config + 3 * (size_t)sv), ld_ag(config + 3 * (size_t)sv + 1),
                         ld_ag(config + 3 * (size_t)sv + 2)};
    double e2[3], uacc = 0.0, z[NLM];
    if (propose_core<LMAX>(seed, sv, sweep, e, step, lane % PROPOSE_LANES, e2, &uacc, z) && ok) {
      double* o = prop + (size_t)qg * PS;
      st_ag(o, e2[0]);
      st_ag(o + 1, e2[1]);
      st_ag(o + 2, e2[2]);
      st_ag(o + 3, uacc);
#pragma unroll
      for (int kk = 0; kk < NLM; ++kk) st_ag(o + 4 + kk, z[kk]);
    }
  }
}

template <int LMAX>
__global__ __launch_bounds__(64) void k_flow(FlowArgs a) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1);
  extern __shared__ double lds[];
  const DevTables& T = a.T;
  const int lane = threadIdx.x;
  double* Z = lds;
  double* dz = Z + T.nb_max * NLM + 1;
  const int n = T.n_active, W = gridDim.x;
  const int nbm = T.nb_max;
  if ((int)blockIdx.x >= n) return;
  // the first item's position tables (each item prefetches the next one's)
  int q = blockIdx.x;
  int s_nx = T.color_sites[q];
  unsigned long long low_nx = a.low[q];
  int nb_nx = lane < nbm ? T.nbq[(size_t)q * nbm + lane] : -1;
  for (int i = 1; i <= a.nsweeps; ++i) {
    const int t = (int)(a.t0 + i);
    const int slot = flow_meas_slot(a.i0, i, a.nsweeps, a.measure_interval, a.defer_last);
    for (; q < n; q += W) {
      const int s = __builtin_amdgcn_readfirstlane(s_nx);
      const unsigned long long low = low_nx;
      const int nb[1] = {nb_nx >= 0 ? nb_nx : s};
      // the next item's position tables, loaded while this one walks
      auto next = [&] {
        int qn = q + W;
        if (qn >= n) qn = blockIdx.x;  // the next sweep's first item
        s_nx = T.color_sites[qn];
        low_nx = a.low[qn];
        nb_nx = lane < nbm ? T.nbq[(size_t)qn * nbm + lane] : -1;
      };
      if (!flow_update<LMAX, 64>(T, a.config, a.zrows, a.prop, a.done, a.err, a.spin_limit, q, s, low, nb, true, t,
                                 a.beta, Z, dz, a.dE_seg + (size_t)(i - 1) * n + q,
                                 a.acc_seg + (size_t)(i - 1) * n + q,
                                 slot >= 0 ? a.cfg_seg + ((size_t)slot * T.n_sites + s) * 3 : nullptr, next))
        return;
    }
    q = blockIdx.x;
    // the next sweep's proposals of this wave's positions, from their spins
    // after this sweep (nothing else writes them)
    if (i < a.nsweeps) flow_propose<LMAX>(T, a.config, a.prop, (int)blockIdx.x, W, a.seed, t + 1, a.step);
  }
}

__global__ void k_flow_init(int* done, int n, int t0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) done[i] = t0;
}
__global__ void k_flow_inactive(int* done, const uint8_t* __restrict__ active, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !active[i]) done[i] = 2147483647;
}

// first reduction level of every sweep of a segment: block (sweep, chunk)
__global__ __launch_bounds__(256) void k_flow_chunks(const double* __restrict__ dE_seg, const int* __restrict__ acc_seg,
                                                     const double* __restrict__ cfg_seg, const uint8_t* __restrict__ active,
                                                     int n, int n_sites, double* part, long long* ipart, long long pstride,
                                                     long long istride, int per_sweep, int nsweeps, long long i0, int mi,
                                                     int defer_last) {
  const int i = blockIdx.x / per_sweep, b = blockIdx.x % per_sweep;
  const int slot = flow_meas_slot(i0, i + 1, nsweeps, mi, defer_last);
  chunk_body<MODE_METRO>(b, dE_seg + (size_t)i * n, n, acc_seg + (size_t)i * n, nullptr, part + i * pstride,
                         ipart + i * istride, slot >= 0 ? cfg_seg + (size_t)slot * 3 * n_sites : nullptr, active,
                         n_sites, slot >= 0 ? 1 : 0);
}

// the segment's final reductions, in sweep order (energy, counters, sweep
// index, measurement rows and the device binner advance as per sweep)
__global__ __launch_bounds__(256) void k_flow_final(double* part, long long* ipart, long long pstride, long long istride,
                                                    int n, int n_sites, ChainDev* st, double* obs, double* scratch,
                                                    DevBinner* bin, int nsweeps, long long i0, int mi, int defer_last) {
  for (int i = 0; i < nsweeps; ++i) {
    const int slot = flow_meas_slot(i0, i + 1, nsweeps, mi, defer_last);
    final_body<MODE_METRO>(part + i * pstride, n, ipart + i * istride, n_sites, slot >= 0 ? 1 : 0, st, obs, n,
                           (double)n, scratch, slot >= 0 ? bin : nullptr);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// renormalisation + full energy (state.jl:168-185, energy.jl:55-75)
// ---------------------------------------------------------------------------
template <int LMAX>
__global__ void k_renorm(DevTables T, double* __restrict__ config, double* __restrict__ zrows) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= T.n_sites || !T.active[s]) return;
  double* e = config + 3 * (size_t)s;
  const double n = __builtin_sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]);
  const double u0 = e[0] / n, u1 = e[1] / n, u2 = e[2] / n;
  e[0] = u0;
  e[1] = u1;
  e[2] = u2;
  double z[NLM];
  sce::zlm_row<LMAX>(u0, u1, u2, z);
#pragma unroll
  for (int k = 0; k < NLM; ++k) zrows[(size_t)s * NLM + k] = z[k];
}

// Tesseral rows of every site from a configuration (3 x n_sites), one lane
// per site (gpu_gradient.jl:41-51: the rows the gradient reads).
template <int LMAX>
__global__ void k_rows_of(int n_sites, const double* __restrict__ config, double* __restrict__ zrows) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sites) return;
  double z[NLM];
  sce::zlm_row<LMAX>(config[3 * (size_t)s], config[3 * (size_t)s + 1], config[3 * (size_t)s + 2], z);
#pragma unroll
  for (int k = 0; k < NLM; ++k) zrows[(size_t)s * NLM + k] = z[k];
}

// The all-site energy gradient G[s] = dE/de_s, tangent-projected — the
// arithmetic contract of the reference's _grad_kernel! / _gradient_lane_ref!
// (gpu/gpu_gradient.jl:92-164, 242-279): one wavefront per site; VL strided
// shares (the reference's workgroup size), share v taking adjacency programs
// j = site_ptr[s] + v, + VL, ... and, per entry with a nonzero product p (the
// sweep's skips), adding (w p) grad Z_tgt(e_s) — w p formed first, then one
// multiply and one add per component; the shares are added in share order.
// Lane k < nlm computes grad Z_k(e_s) into LDS (each element is independent of
// the others).  Writes G[s] only; reads config and rows.
template <int LMAX, int VL>
__global__ __launch_bounds__(TPB) void k_grad(GradTables T, const double* __restrict__ config,
                                              const double* __restrict__ zrows, double* __restrict__ G) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1), W = TPB / 64, K = VL / 64;
  __shared__ double grow[W][3 * NLM];
  __shared__ double part[W][3 * VL];
  const int w = threadIdx.x / 64, h = threadIdx.x % 64;
  const int s = blockIdx.x * W + w;  // 0-based site
  const bool valid = s < T.n_sites;
  if (valid && h < NLM) {
    int l, m;
    sce::lm_of(h, &l, &m);
    double g[3];
    sce::grad_zlm_one<LMAX>(l, m, config[3 * (size_t)s], config[3 * (size_t)s + 1], config[3 * (size_t)s + 2], g);
    grow[w][3 * h] = g[0];
    grow[w][3 * h + 1] = g[1];
    grow[w][3 * h + 2] = g[2];
  }
  __syncthreads();
  if (valid) {
    const int j0 = T.site_ptr[s], j1 = T.site_ptr[s + 1];  // 1-based [j0, j1)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int v = h + 64 * k;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0;
      for (int j = j0 + v; j < j1; j += VL) {
        const int pid = T.site_prog[j - 1], col = T.site_col[j - 1];
        const int e0 = T.sprog_ptr[pid - 1], e1 = T.sprog_ptr[pid];
        const int col2 = col < 0 ? T.site_col2[j - 1] : 0;
        // a general program's instance members (1-based sites), read once
        // (bodies have at most 4 members: the table upload's layout check)
        int mem[4] = {1, 1, 1, 1};
        if (col == 0) {
          const int inst = T.site_inst[j - 1];
          const int off = T.inst_ptr[inst - 1] - 1, body = T.inst_ptr[inst] - 1 - off;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (q < body) mem[q] = T.inst_sites[off + q];
        }
        for (int e = e0; e < e1; ++e) {
          double p;
          if (col > 0) {
            p = zrows[(size_t)(col - 1) * T.nlm + (T.pent_row[e - 1] - 1)];
          } else if (col < 0) {
            p = zrows[(size_t)(-col - 1) * T.nlm + (T.pent_row[e - 1] - 1)] *
                zrows[(size_t)(col2 - 1) * T.nlm + (T.pent_row2[e - 1] - 1)];
          } else {
            p = 1.0;
            for (int f = T.sfac_ptr[e - 1]; f < T.sfac_ptr[e]; ++f) {
              const int sl = T.sfac_slot[f - 1];
              const int mm = sl == 1 ? mem[0] : sl == 2 ? mem[1] : sl == 3 ? mem[2] : mem[3];
              p *= zrows[(size_t)(mm - 1) * T.nlm + (T.sfac_row[f - 1] - 1)];
            }
          }
          if (p == 0.0) continue;
          const int t = T.sent_tgt[e - 1] - 1;
          const double wp = T.sent_w[e - 1] * p;
          a0 = a0 + wp * grow[w][3 * t];
          a1 = a1 + wp * grow[w][3 * t + 1];
          a2 = a2 + wp * grow[w][3 * t + 2];
        }
      }
      part[w][3 * v] = a0;
      part[w][3 * v + 1] = a1;
      part[w][3 * v + 2] = a2;
    }
  }
  __syncthreads();
  if (valid && h < 3) {
    // component h of the share-ordered fold
    double g = 0.0;
    for (int v = 0; v < VL; ++v) g += part[w][3 * v + h];
    G[3 * (size_t)s + h] = g;
  }
}

// The same gradient on the walk layout: one wavefront per active site (colour
// position q), the neighbour rows staged in LDS as the walk stages them, and
// each raw adjacency entry's program read from the sublattice's pid-sorted
// walk (adj_perm maps raw to sorted index; the walk entry's product
// (Z[f0] Z[f1]) Z[f2] with unused factors 1.0 is the reference's p of every
// path bit for bit).  A program's entries are loaded four at a time (clamped
// indices, the extra terms dropped by selects) and added in entry order.
// Shares, entry order, skips and fold as k_grad.  Rows of inactive sites are
// not written (the caller zeroes G).
template <int LMAX, int VL>
__global__ __launch_bounds__(TPB) void k_grad_walk(DevTables T, const double* __restrict__ config,
                                                   const double* __restrict__ zrows, double* __restrict__ G) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1), W = TPB / 64, K = VL / 64;
  extern __shared__ double lds[];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), h = threadIdx.x % 64;
  const int zs = T.nb_max * NLM + 1;
  double* Z = lds + w * (zs + 3 * NLM + 3 * VL);
  double* grow = Z + zs;
  double* part = grow + 3 * NLM;
  const int q = __builtin_amdgcn_readfirstlane((int)blockIdx.x * W + w);
  const bool valid = q < T.n_active;
  int s = 0, a = 0;
  if (valid) {
    s = __builtin_amdgcn_readfirstlane(T.color_sites[q]);
    a = stage_neighbours_q<64, NLM>(T, zrows, q, s, h, Z, [] {});
    if (h < NLM) {
      int l, m;
      sce::lm_of(h, &l, &m);
      double g[3];
      sce::grad_zlm_one<LMAX>(l, m, config[3 * (size_t)s], config[3 * (size_t)s + 1], config[3 * (size_t)s + 2], g);
      grow[3 * h] = g[0];
      grow[3 * h + 1] = g[1];
      grow[3 * h + 2] = g[2];
    }
  }
  __syncthreads();
  if (valid) {
    const int pp = T.prog_ptr[a], nadj = T.prog_ptr[a + 1] - pp - 1;
    const WalkEntry* __restrict__ wl = T.walk + T.walk_ptr[a];
    const uint16_t* __restrict__ perm = T.adj_perm + (size_t)q * T.adj_max;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int v = h + 64 * k;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0;
      for (int r = v; r < nadj; r += VL) {
        const int kk = perm[r];
        const int e1 = T.prog_off[pp + kk + 1];
        for (int base = T.prog_off[pp + kk]; base < e1; base += 4) {
          WalkEntry en[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) en[u] = wl[base + u < e1 ? base + u : e1 - 1];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const double p = (Z[en[u].f0] * Z[en[u].f1]) * Z[en[u].f2];
            const bool use = base + u < e1 && p != 0.0;
            const double wp = en[u].w * p;
            const double b0 = a0 + wp * grow[3 * en[u].tgt];
            const double b1 = a1 + wp * grow[3 * en[u].tgt + 1];
            const double b2 = a2 + wp * grow[3 * en[u].tgt + 2];
            a0 = use ? b0 : a0;
            a1 = use ? b1 : a1;
            a2 = use ? b2 : a2;
          }
        }
      }
      part[3 * v] = a0;
      part[3 * v + 1] = a1;
      part[3 * v + 2] = a2;
    }
  }
  __syncthreads();
  if (valid && h < 3) {
    double g = 0.0;
    for (int v = 0; v < VL; ++v) g += part[3 * v + h];
    G[3 * (size_t)s + h] = g;
  }
}

__global__ void k_inst_energy(DevTables T, const double* __restrict__ zrows,
                              double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.n_inst) return;
  const int k = T.inst_term[i];
  const int off = T.inst_ptr[i];
  double Ei = 0.0;
  for (int e = T.eprog_ptr[k]; e < T.eprog_ptr[k + 1]; ++e) {
    double p = 1.0;
    const int f0 = T.efac_ptr[e];
    const int nf = T.efac_ptr[e + 1] - f0;
    for (int m = 0; m < nf; ++m)
      p *= zrows[(size_t)T.inst_sites[off + m] * T.nlm + T.efac_row[f0 + m]];
    Ei += T.eent_w[e] * p;
  }
  out[i] = T.term_coef[k] * Ei;
}

__global__ void k_set_params(ChainDev* st, double beta, double step, int set_step) {
  st->beta = beta;
  if (set_step) st->step = step;
}

// a call's measurement rows start at row 0 (in stream order, no host round trip)
__global__ void k_reset_nobs(ChainDev* st) { st->n_obs = 0; }

// ---------------------------------------------------------------------------
// diagnostics kernels
// ---------------------------------------------------------------------------
__global__ void k_test_philox(long long n, const uint64_t* seeds, const uint32_t* ctr,
                              uint32_t* out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  sce::u32x4 c{ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]};
  const sce::u32x4 r = sce::philox4x32(c, (uint32_t)seeds[i], (uint32_t)(seeds[i] >> 32));
  out[4 * i] = r.x;
  out[4 * i + 1] = r.y;
  out[4 * i + 2] = r.z;
  out[4 * i + 3] = r.w;
}

__global__ void k_test_detmath(long long n, int op, const double* x, double* out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s, c;
  switch (op) {
    case 0: out[i] = sce::det_exp(x[i]); break;
    case 1: out[i] = sce::det_log(x[i]); break;
    case 2: sce::det_sincos(x[i], &s, &c); out[i] = s; break;
    case 4: out[i] = (double)sce::det_tanhf((float)x[i]); break;
    default: sce::det_sincos(x[i], &s, &c); out[i] = c; break;
  }
}

template <int LMAX>
__global__ void k_test_zlm(long long n, const double* dirs, double* out) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1);
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double z[NLM];
  sce::zlm_row<LMAX>(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2], z);
  for (int k = 0; k < NLM; ++k) out[i * NLM + k] = z[k];
}

__global__ void k_test_proposal(long long n, uint64_t seed, const int32_t* sites,
                                const int32_t* sweeps, const double* e3, double step,
                                double* e2o, double* uo) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double e[3] = {e3[3 * i], e3[3 * i + 1], e3[3 * i + 2]};
  double e2[3], u;
  sce::keyed_proposal(seed, sites[i], sweeps[i], e, step, e2, &u);
  e2o[3 * i] = e2[0];
  e2o[3 * i + 1] = e2[1];
  e2o[3 * i + 2] = e2[2];
  uo[i] = u;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Owner of device allocations: everything dalloc'd into it is freed with it;
// release() frees one allocation early (a buffer replaced by a larger one).
struct DevBuf {
  std::vector<void*> ptrs;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  void release(void* p) {
    if (!p) return;
    for (size_t i = 0; i < ptrs.size(); ++i)
      if (ptrs[i] == p) {
        (void)hipFree(p);
        ptrs[i] = ptrs.back();
        ptrs.pop_back();
        return;
      }
  }
  bool owns(const void* p) const {
    for (void* q : ptrs)
      if (q == p) return true;
    return false;
  }
  // Hand allocation `mine` to `other` and take `theirs` from it (a payload swap
  // moves ownership with the pointers, so each owner frees what it holds).
  static void exchange(DevBuf& a, void* mine, DevBuf& b, void* theirs) {
    for (void*& q : a.ptrs)
      if (q == mine) {
        q = theirs;
        break;
      }
    for (void*& q : b.ptrs)
      if (q == theirs) {
        q = mine;
        break;
      }
  }
};

// A captured graph and its executable.  The source graph lives as long as the
// executable (tools that inspect a launched executable may walk its source
// nodes), and both are destroyed only once the stream replaying them has
// drained.
struct GraphSlot {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  explicit operator bool() const { return x != nullptr; }
  void reset() {
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
    x = nullptr;
    g = nullptr;
  }
};

// Capture what `enqueue` issues on stream s into slot (empty on entry).
template <class F>
static int32_t capture_graph(hipStream_t s, GraphSlot& slot, F&& enqueue) {
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return sce_fail(SCE_EHIP, "hipStreamBeginCapture failed");
  const int32_t r = enqueue();
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (r != SCE_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return r != SCE_OK ? r : sce_fail(SCE_EHIP, "stream capture failed: %s", hipGetErrorString(e));
  }
  e = hipGraphInstantiate(&slot.x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    slot.x = nullptr;
    (void)hipGraphDestroy(g);
    return sce_fail(SCE_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  slot.g = g;
  return SCE_OK;
}

template <class T>
static int32_t dalloc(DevBuf& b, T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess)
    return sce_fail(e == hipErrorOutOfMemory ? SCE_ENOMEM : SCE_EHIP, "hipMalloc(%zu): %s",
                    n * sizeof(T), hipGetErrorString(e));
  b.ptrs.push_back(*p);
  return SCE_OK;
}

template <class T>
static int32_t dupload(DevBuf& b, T** p, const T* src, size_t n) {
  CHK(dalloc(b, p, n));
  if (n) HIPCHK(hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice));
  return SCE_OK;
}

// As dupload, into an allocation of at least one element (a valid pointer for
// empty tables).
template <class T>
static int32_t dupload1(DevBuf& b, T** p, const T* src, size_t n) {
  CHK(dalloc(b, p, std::max<size_t>(n, 1)));
  if (n) HIPCHK(hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice));
  return SCE_OK;
}

// Rows of a measurement buffer that must hold `want`: at least 4096 and at
// least twice the old capacity, so that captured sweeps (which hold the
// buffer's address) are rarely dropped for a new one.
static size_t obs_rows_for(size_t want, size_t old_cap) {
  return std::max(want, std::max<size_t>(4096, 2 * old_cap));
}

struct sce_tables {
  DevBuf mem;
  DevTables dev{};
  std::vector<int32_t> color_ptr;  // 0-based offsets, n_colors + 1
  std::vector<int32_t> color_sites;  // 0-based site ids, colour order (host copy)
  int32_t* d_color_ptr = nullptr;
  int max_class = 0;
  int fin_col = -1;  // the smallest non-empty colour class: its launch carries a deferred final
  std::vector<int> sub_F;  // surrogate descriptor width per sublattice (0: inactive)
  std::vector<std::vector<float>> sub_qin;  // int8 input scale of each descriptor feature
  int n_colors = 0;
  int64_t total_walk = 0;
  int w_max = 0;
  GradTables grad{};
};

struct sce_chain {
  sce_tables* t = nullptr;
  DevBuf mem;
  double *config = nullptr, *zrows = nullptr, *prop = nullptr, *dE = nullptr, *partial = nullptr,
         *inst_e = nullptr, *obs = nullptr;
  int *acc = nullptr, *att = nullptr, *ipart = nullptr, *err = nullptr;
  ChainDev* st = nullptr;
  // dependency-driven persistent sweeps: segment buffers, sized on first use
  struct Flow {
    int cap = 0, mcap = 0;  // sweeps and measured sweeps the buffers hold
    double *dE = nullptr, *cfg = nullptr, *part = nullptr, *scratch = nullptr;
    int *acc = nullptr, *done = nullptr;
    long long* ipart = nullptr;
  } flow;
  long long* ip64 = nullptr;  // integer partials of the reductions
  bool persistent = false;  // run sweeps in one cooperative launch (opt-in)
  bool use_graph = true;    // replay captured sweeps as hipGraphs
  DevBinner* bin = nullptr;  // device binner fed by every measurement (null: off)
  DevBinner* bin_mem = nullptr;
  // captured sweeps: [fused measurement + 2 need_prop + 4 next_prop (with the final
  // deferred) + 8 measurement flag of the previous sweep's deferred final]
  GraphSlot gexec[16];
  int g_lanes = 0;
  // captured compound sweeps (Metropolis + overrelaxation passes):
  // [fused measurement + 2 need_prop + 4 next_prop], for cg_or passes
  GraphSlot cgexec[8];
  int cg_or = 0;
  size_t obs_cap = 0;
  hipStream_t stream = nullptr;
  long long sweeps_done = 0;  // host mirror of sweep_index
  double beta_dev = std::nan("");
  // timing of the last sce_run_sweeps
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  bool timing = false;
  bool timing_enabled = false;
  double total_ms = 0.0, walk_ms = 0.0;
  long long walk_launches = 0;
  void drop_graphs() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& g : gexec) g.reset();
    for (auto& g : cgexec) g.reset();
    g_lanes = 0;
    cg_or = 0;
  }
  ~sce_chain() {
    drop_graphs();
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static size_t partial_capacity(long long n) {
  size_t cap = 16;
  long long m = (n + TREE_B - 1) / TREE_B;
  while (m > 1) {
    cap += (size_t)m + 2;
    m = (m + TREE_B - 1) / TREE_B;
  }
  return 3 * cap + 16;
}

// The reference's adjacency programs for the energy gradient, uploaded as the
// descriptor holds them (1-based).
// ---------------------------------------------------------------------------
// Device-side table build: the per-position tables of the walk (neighbour
// sites, lower-colour masks, raw -> pid-sorted adjacency permutation) made on
// the device from the uploaded reference tables, each site verified against
// its sublattice's representative adjacency translated to its cell (the
// host's sce_derive_layout check, one wavefront per position).
// ---------------------------------------------------------------------------
struct LayoutBuild {
  int n_sites, n_active, na, d0, d1, d2, nb_max, adj_max, n_colors;
  const int32_t* color_sites;  // 0-based
  const int32_t* color_ptr;    // 0-based offsets
  const uint8_t* active;
  const int32_t* nb_ptr;
  const int4* nb;
  // the reference tables, 1-based as uploaded
  const int32_t *site_ptr, *site_prog, *site_col, *site_col2, *site_inst, *inst_ptr, *inst_sites;
  const int8_t* site_slot;
  // the representative adjacency per sublattice
  const int32_t* rep_ptr;
  const int32_t* rep_pid;
  const int8_t* rep_cls;
  const int4* rep_nb;
  int32_t* nbq;
  unsigned long long* low;
  uint16_t* perm;
  int* col_of;
  unsigned long long* err;  // min over failures of (site + 1) * 8 + reason
};

__device__ __forceinline__ int lay_translate(const LayoutBuild& B, int a, int k, int c1, int c2, int c3) {
  const int4 o = B.nb[B.nb_ptr[a] + k];
  return o.x + B.na * ((c1 + o.y) % B.d0 + B.d0 * ((c2 + o.z) % B.d1 + B.d1 * ((c3 + o.w) % B.d2)));
}
__device__ __forceinline__ void lay_fail(const LayoutBuild& B, int s, int reason) {
  atomicMin(B.err, (unsigned long long)(s + 1) * 8ull + (unsigned long long)reason);
}

__global__ void k_lay_sites(LayoutBuild B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.n_sites) return;
  if ((B.active[s] != 0) != (B.active[s % B.na] != 0)) lay_fail(B, s, 1);  // activity differs
}

__global__ void k_lay_colour(LayoutBuild B) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B.n_active) return;
  int c = 0;
  while (c + 1 < B.n_colors && B.color_ptr[c + 1] <= q) ++c;
  B.col_of[B.color_sites[q]] = c;
}

__global__ void k_lay_nbq(LayoutBuild B) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B.n_active) return;
  const int s = B.color_sites[q], a = s % B.na, cell = s / B.na;
  const int c1 = cell % B.d0, c2 = (cell / B.d0) % B.d1, c3 = cell / (B.d0 * B.d1);
  const int cnt = B.nb_ptr[a + 1] - B.nb_ptr[a];
  for (int j = 0; j < B.nb_max; ++j) B.nbq[(size_t)q * B.nb_max + j] = j < cnt ? lay_translate(B, a, j, c1, c2, c3) : s;
}

__global__ void k_lay_low(LayoutBuild B) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B.n_active) return;
  const int cq = B.col_of[B.color_sites[q]];
  unsigned long long m = 0ull;
  for (int j = 0; j < B.nb_max; ++j) {
    const int cn = B.col_of[B.nbq[(size_t)q * B.nb_max + j]];
    if (cn >= 0 && cn < cq) m |= 1ull << j;
  }
  B.low[q] = m;
}

// one wavefront per colour position
__global__ __launch_bounds__(64) void k_lay_verify(LayoutBuild B) {
  const int q = blockIdx.x;
  if (q >= B.n_active) return;
  const int lane = threadIdx.x;
  const int s = B.color_sites[q], a = s % B.na, cell = s / B.na;
  const int c1 = cell % B.d0, c2 = (cell / B.d0) % B.d1, c3 = cell / (B.d0 * B.d1);
  const int j0 = B.site_ptr[s], n = B.site_ptr[s + 1] - j0;
  const int r0 = B.rep_ptr[a];
  if (n != B.rep_ptr[a + 1] - r0) {
    if (lane == 0) lay_fail(B, s, 2);  // adjacency length
    return;
  }
  for (int k = lane; k < n; k += 64) {
    const int j = j0 + k;  // 1-based adjacency entry
    const int pid = B.site_prog[j - 1];
    // stable rank by program id
    int rank = 0;
    for (int m = 0; m < n; ++m) {
      const int pm = B.site_prog[j0 + m - 1];
      rank += (pm < pid || (pm == pid && m < k)) ? 1 : 0;
    }
    B.perm[(size_t)q * B.adj_max + k] = (uint16_t)rank;
    const int r = r0 + rank;
    const int col = B.site_col[j - 1];
    const int cls = col > 0 ? 1 : col < 0 ? 2 : 0;
    if (B.rep_pid[r] != pid || B.rep_cls[r] != cls) {
      lay_fail(B, s, 3);  // program sequence
      continue;
    }
    int fs[3], nf = 0;
    if (col > 0) {
      fs[nf++] = col;
    } else if (col < 0) {
      fs[nf++] = -col;
      fs[nf++] = B.site_col2[j - 1];
    } else {
      const int inst = B.site_inst[j - 1];
      const int off = B.inst_ptr[inst - 1] - 1;
      const int body = B.inst_ptr[inst] - B.inst_ptr[inst - 1];
      const int own = B.site_slot[j - 1];
      if (body - 1 > 3) {
        lay_fail(B, s, 4);
        continue;
      }
      for (int x = 1; x <= body; ++x)
        if (x != own) fs[nf++] = B.inst_sites[off + x - 1];
    }
    const int4 rn = B.rep_nb[r];
    const int want[3] = {rn.y, rn.z, rn.w};
    if (nf != rn.x) {
      lay_fail(B, s, 4);  // neighbour / factor count
      continue;
    }
    for (int f = 0; f < nf; ++f)
      if (lay_translate(B, a, want[f], c1, c2, c3) + 1 != fs[f]) lay_fail(B, s, 4);
  }
}

static int32_t upload_grad_tables(sce_tables* t, const sce_tables_desc* d) {
  GradTables& g = t->grad;
  g.n_sites = d->n_sites;
  g.nlm = d->nlm;
  const size_t na = (size_t)d->n_inst_sites, np = (size_t)d->n_prog, ns = (size_t)d->n_sent;
  size_t n_inst_members = 0;
  if (d->n_inst > 0) n_inst_members = (size_t)(d->inst_ptr[d->n_inst] - 1);
  int32_t* q;
  int8_t* q8;
  double* qd;
  CHK(dupload(t->mem, &q, d->site_ptr, (size_t)d->n_sites + 1)); g.site_ptr = q;
  CHK(dupload1(t->mem, &q, d->site_inst, na)); g.site_inst = q;
  CHK(dupload1(t->mem, &q, d->site_prog, na)); g.site_prog = q;
  CHK(dupload1(t->mem, &q, d->site_col, na)); g.site_col = q;
  CHK(dupload1(t->mem, &q, d->site_col2, na)); g.site_col2 = q;
  CHK(dupload(t->mem, &q, d->sprog_ptr, np + 1)); g.sprog_ptr = q;
  CHK(dupload1(t->mem, &q, d->sent_tgt, ns)); g.sent_tgt = q;
  CHK(dupload(t->mem, &q, d->sfac_ptr, ns + 1)); g.sfac_ptr = q;
  CHK(dupload1(t->mem, &q, d->sfac_row, (size_t)d->n_sfac)); g.sfac_row = q;
  CHK(dupload1(t->mem, &q8, d->sfac_slot, (size_t)d->n_sfac)); g.sfac_slot = q8;
  CHK(dupload1(t->mem, &q, d->pent_row, ns)); g.pent_row = q;
  CHK(dupload1(t->mem, &q, d->pent_row2, ns)); g.pent_row2 = q;
  CHK(dupload(t->mem, &q, d->inst_ptr, (size_t)d->n_inst + 1)); g.inst_ptr = q;
  CHK(dupload1(t->mem, &q, d->inst_sites, n_inst_members)); g.inst_sites = q;
  CHK(dupload1(t->mem, &qd, d->sent_w, ns)); g.sent_w = qd;
  return SCE_OK;
}

// The per-position tables of the walk, made on the device (k_lay_*), with every
// site verified against its sublattice's representative adjacency; the
// temporaries are freed again.
static int32_t build_position_tables(sce_tables* t, const sce_tables_desc* d, const HostLayout& L) {
  DevTables& D = t->dev;
  const GradTables& G = t->grad;
  LayoutBuild B{};
  B.n_sites = d->n_sites;
  B.n_active = d->n_active;
  B.na = L.na;
  B.d0 = d->dims[0];
  B.d1 = d->dims[1];
  B.d2 = d->dims[2];
  B.nb_max = L.nb_max;
  B.adj_max = D.adj_max;
  B.n_colors = d->n_colors;
  B.color_sites = D.color_sites;
  B.color_ptr = t->d_color_ptr;
  B.active = D.active;
  B.nb_ptr = D.nb_ptr;
  B.nb = D.nb;
  B.site_ptr = G.site_ptr;
  B.site_prog = G.site_prog;
  B.site_col = G.site_col;
  B.site_col2 = G.site_col2;
  B.site_inst = G.site_inst;
  B.inst_ptr = G.inst_ptr;
  B.inst_sites = G.inst_sites;
  const size_t nact = (size_t)std::max(d->n_active, 1);
  std::vector<void*> tmp;
  auto release = [&] {
    for (void* p : tmp) t->mem.release(p);
  };
  auto up_tmp = [&](auto** p, const auto* src, size_t n) -> int32_t {
    int32_t r = dupload1(t->mem, p, src, n);
    if (r == SCE_OK) tmp.push_back((void*)*p);
    return r;
  };
  int32_t r = SCE_OK;
  int8_t* slot;
  int32_t *rp, *rpid;
  int8_t* rcls;
  int* col_of;
  unsigned long long* err;
  if ((r = up_tmp(&slot, d->site_slot, (size_t)d->n_inst_sites)) != SCE_OK ||
      (r = up_tmp(&rp, L.rep_ptr.data(), L.rep_ptr.size())) != SCE_OK ||
      (r = up_tmp(&rpid, L.rep_pid.data(), L.rep_pid.size())) != SCE_OK ||
      (r = up_tmp(&rcls, L.rep_cls.data(), L.rep_cls.size())) != SCE_OK) {
    release();
    return r;
  }
  int4* rnb;
  if ((r = up_tmp(&rnb, reinterpret_cast<const int4*>(L.rep_nb.data()), L.rep_nb.size() / 4)) != SCE_OK ||
      (r = dalloc(t->mem, &col_of, (size_t)std::max(d->n_sites, 1))) != SCE_OK) {
    release();
    return r;
  }
  tmp.push_back(col_of);
  if ((r = dalloc(t->mem, &err, 1)) != SCE_OK) {
    release();
    return r;
  }
  tmp.push_back(err);
  B.site_slot = slot;
  B.rep_ptr = rp;
  B.rep_pid = rpid;
  B.rep_cls = rcls;
  B.rep_nb = rnb;
  B.col_of = col_of;
  B.err = err;
  // the tables that stay: neighbour sites, lower-colour masks, permutation
  int32_t* nbq;
  uint16_t* perm;
  if ((r = dalloc(t->mem, &nbq, nact * std::max(L.nb_max, 1))) != SCE_OK ||
      (r = dalloc(t->mem, &perm, nact * D.adj_max)) != SCE_OK) {
    release();
    return r;
  }
  B.nbq = nbq;
  B.perm = perm;
  if (L.nb_max <= 64) {
    unsigned long long* low;
    if ((r = dalloc(t->mem, &low, nact)) != SCE_OK) {
      release();
      return r;
    }
    B.low = low;
    D.nbq_low = low;
  }
  D.nbq = nbq;
  D.adj_perm = perm;
  const unsigned long long none = ~0ull;
  hipError_t e = hipMemcpy(err, &none, sizeof none, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(col_of, 0xff, (size_t)std::max(d->n_sites, 1) * sizeof(int));
  if (e == hipSuccess) e = hipMemset(perm, 0, nact * D.adj_max * sizeof(uint16_t));
  if (e != hipSuccess) {
    release();
    return sce_fail(SCE_EHIP, "layout build: %s", hipGetErrorString(e));
  }
  const int N = d->n_sites, n = d->n_active;
  if (N > 0) k_lay_sites<<<(N + 255) / 256, 256>>>(B);
  if (n > 0) {
    k_lay_colour<<<(n + 255) / 256, 256>>>(B);
    k_lay_nbq<<<(n + 255) / 256, 256>>>(B);
    if (B.low) k_lay_low<<<(n + 255) / 256, 256>>>(B);
    k_lay_verify<<<n, 64>>>(B);
  }
  unsigned long long code = none;
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(&code, err, sizeof code, hipMemcpyDeviceToHost);
  release();
  if (e != hipSuccess) return sce_fail(SCE_EHIP, "layout build: %s", hipGetErrorString(e));
  if (code != none) {
    static const char* why[] = {"", "activity differs from its sublattice", "adjacency length differs from its sublattice",
                                "program sequence differs from its sublattice",
                                "a neighbour is not the translated stencil"};
    const int site = (int)(code / 8), reason = (int)(code % 8);
    return sce_fail(SCE_ELAYOUT, "site %d: %s", site, why[reason < 5 ? reason : 0]);
  }
  return SCE_OK;
}

extern "C" int32_t sce_tables_create(const sce_tables_desc* d, sce_tables** out) {
  if (!d || !out) return sce_fail(SCE_EINVAL, "sce_tables_create: NULL argument");
  *out = nullptr;
  if (d->lmax < 0 || d->lmax > 6)
    return sce_fail(SCE_EINVAL, "lmax = %d unsupported on the device path (<= 6)", d->lmax);
  if (d->nlm != (d->lmax + 1) * (d->lmax + 1)) return sce_fail(SCE_EDIM, "nlm != (lmax+1)^2");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return sce_fail(SCE_ENODEV, "no HIP device available");
  HostLayout L;
  CHK(sce_derive_layout(d, &L, false));  // every site is verified by the device build below
  auto* t = new sce_tables();
  auto fail = [&](int32_t r) {
    delete t;
    return r;
  };
  DevTables& D = t->dev;
  D.n_sites = d->n_sites;
  D.n_active = d->n_active;
  D.nlm = d->nlm;
  D.lmax = d->lmax;
  D.na = d->n_cell_atoms;
  D.d0 = d->dims[0];
  D.d1 = d->dims[1];
  D.d2 = d->dims[2];
  D.nb_max = L.nb_max;
  D.n_inst = d->n_inst;
  std::vector<int32_t> cs(d->n_active);
  for (int q = 0; q < d->n_active; ++q) cs[q] = d->color_sites[q] - 1;
  int32_t r;
  int32_t* p32;
  if ((r = dupload(t->mem, &p32, cs.data(), cs.size())) != SCE_OK) return fail(r);
  D.color_sites = p32;
  t->color_sites = cs;
  uint8_t* p8;
  if ((r = dupload(t->mem, &p8, d->site_has_l1, (size_t)d->n_sites)) != SCE_OK) return fail(r);
  D.has_l1 = p8;
  if ((r = dupload(t->mem, &p8, d->site_active, (size_t)d->n_sites)) != SCE_OK) return fail(r);
  D.active = p8;
  if ((r = dupload(t->mem, &p32, L.walk_ptr.data(), L.walk_ptr.size())) != SCE_OK) return fail(r);
  D.walk_ptr = p32;
  WalkEntry* pw;
  if ((r = dupload(t->mem, &pw, L.walk.data(), L.walk.size())) != SCE_OK) return fail(r);
  D.walk = pw;
  {
    // the padded copy the walks read, in LDS byte offsets from the staged
    // block: each sublattice's entries, then no-op entries (weight 0, factors
    // on its 1.0 slot, target row 0) up to a multiple of WALK_PAD.  A block
    // beyond 16-bit offsets (64 KB) cannot be staged by any walk launch
    // (check_lds), which then finds no padded walk.
    const size_t dz0 = (size_t)L.nb_max * d->nlm + 1;
    if ((dz0 + d->nlm) * sizeof(double) <= 65535) {
      std::vector<int32_t> pp((size_t)L.na + 1, 0);
      std::vector<int2> pre((size_t)L.na);
      std::vector<WalkEntry> wp;
      auto bytes = [&](const WalkEntry& e) {
        return WalkEntry{e.w, (uint16_t)(8 * e.f0), (uint16_t)(8 * e.f1), (uint16_t)(8 * e.f2),
                         (uint16_t)(8 * (dz0 + e.tgt))};
      };
      for (int a = 0; a < L.na; ++a) {
        pp[a] = (int32_t)wp.size();
        for (int i = L.walk_ptr[a]; i < L.walk_ptr[a + 1]; ++i) wp.push_back(bytes(L.walk[i]));
        const uint16_t one = (uint16_t)((L.nb_ptr[a + 1] - L.nb_ptr[a]) * d->nlm);
        while (wp.size() % WALK_PAD) wp.push_back(bytes(WalkEntry{0.0, one, one, one, 0}));
        int n1 = 0, n2 = 0;
        const int n = (int)wp.size() - pp[a];
        const uint16_t one8 = (uint16_t)(8 * one);
        while (n1 < n && wp[pp[a] + n1].f1 == one8 && wp[pp[a] + n1].f2 == one8) ++n1;
        n2 = n1;
        while (n2 < n && wp[pp[a] + n2].f2 == one8) ++n2;
        pre[a] = make_int2(n1, n2);
      }
      pp[L.na] = (int32_t)wp.size();
      int2* pq;
      if ((r = dupload(t->mem, &pq, pre.data(), pre.size())) != SCE_OK) return fail(r);
      D.walkp_pre = pq;
      if ((r = dupload(t->mem, &p32, pp.data(), pp.size())) != SCE_OK) return fail(r);
      D.walkp_ptr = p32;
      if ((r = dupload(t->mem, &pw, wp.data(), wp.size())) != SCE_OK) return fail(r);
      D.walkp = pw;
    }
  }
  if ((r = dupload(t->mem, &p32, L.nb_ptr.data(), L.nb_ptr.size())) != SCE_OK) return fail(r);
  D.nb_ptr = p32;
  int4* pn;
  if ((r = dupload(t->mem, &pn, reinterpret_cast<const int4*>(L.nb.data()), L.nb.size() / 4)) != SCE_OK)
    return fail(r);
  D.nb = pn;
  t->n_colors = d->n_colors;
  t->color_ptr.resize((size_t)d->n_colors + 1);
  for (int c = 0; c <= d->n_colors; ++c) t->color_ptr[c] = d->color_ptr[c] - 1;
  if ((r = dupload(t->mem, &t->d_color_ptr, t->color_ptr.data(), t->color_ptr.size())) != SCE_OK)
    return fail(r);
  if ((r = upload_grad_tables(t, d)) != SCE_OK) return fail(r);
  {
    // the per-position tables, built and verified on the device
    int amax = 0;
    for (int a = 0; a < L.na; ++a) amax = std::max(amax, L.prog_ptr[a + 1] - L.prog_ptr[a] - 1);
    if (amax > 65535) return fail(sce_fail(SCE_EINVAL, "%d adjacency entries per site exceed the gradient's index", amax));
    D.adj_max = std::max(amax, 1);
    if ((r = build_position_tables(t, d, L)) != SCE_OK) return fail(r);
  }
  {
    if ((r = dupload(t->mem, &p32, L.prog_ptr.data(), L.prog_ptr.size())) != SCE_OK) return fail(r);
    D.prog_ptr = p32;
    if ((r = dupload1(t->mem, &p32, L.prog_off.data(), L.prog_off.size())) != SCE_OK) return fail(r);
    D.prog_off = p32;
  }
  if ((r = dupload(t->mem, &p32, L.feat_ptr.data(), L.feat_ptr.size())) != SCE_OK) return fail(r);
  D.feat_ptr = p32;
  if ((r = dupload(t->mem, &pn, reinterpret_cast<const int4*>(L.feat.data()), L.feat.size() / 4)) != SCE_OK)
    return fail(r);
  D.feat = pn;
  D.nfeat_max = L.nfeat_max;
  // energy tables, 0-based
  std::vector<int32_t> v;
  v.resize(d->n_inst);
  for (int i = 0; i < d->n_inst; ++i) v[i] = d->inst_term[i] - 1;
  if ((r = dupload(t->mem, &p32, v.data(), v.size())) != SCE_OK) return fail(r);
  D.inst_term = p32;
  v.resize((size_t)d->n_inst + 1);
  for (int i = 0; i <= d->n_inst; ++i) v[i] = d->inst_ptr[i] - 1;
  if ((r = dupload(t->mem, &p32, v.data(), v.size())) != SCE_OK) return fail(r);
  D.inst_ptr = p32;
  v.resize((size_t)d->n_inst_sites);
  for (int64_t i = 0; i < d->n_inst_sites; ++i) v[i] = d->inst_sites[i] - 1;
  if ((r = dupload(t->mem, &p32, v.data(), v.size())) != SCE_OK) return fail(r);
  D.inst_sites = p32;
  v.resize((size_t)d->n_terms + 1);
  for (int i = 0; i <= d->n_terms; ++i) v[i] = d->eprog_ptr[i] - 1;
  if ((r = dupload(t->mem, &p32, v.data(), v.size())) != SCE_OK) return fail(r);
  D.eprog_ptr = p32;
  double* pd;
  if ((r = dupload(t->mem, &pd, d->eent_w, (size_t)d->n_eent)) != SCE_OK) return fail(r);
  D.eent_w = pd;
  v.resize((size_t)d->n_eent + 1);
  for (int64_t i = 0; i <= d->n_eent; ++i) v[i] = d->efac_ptr[i] - 1;
  if ((r = dupload(t->mem, &p32, v.data(), v.size())) != SCE_OK) return fail(r);
  D.efac_ptr = p32;
  v.resize((size_t)d->n_efac);
  for (int64_t i = 0; i < d->n_efac; ++i) v[i] = d->efac_row[i] - 1;
  if ((r = dupload(t->mem, &p32, v.data(), v.size())) != SCE_OK) return fail(r);
  D.efac_row = p32;
  if ((r = dupload(t->mem, &pd, d->term_coef, (size_t)d->n_terms)) != SCE_OK) return fail(r);
  D.term_coef = pd;
  for (int c = 0; c < d->n_colors; ++c) {
    const int n = t->color_ptr[c + 1] - t->color_ptr[c];
    t->max_class = std::max(t->max_class, n);
    if (n > 0 && (t->fin_col < 0 || n < t->color_ptr[t->fin_col + 1] - t->color_ptr[t->fin_col])) t->fin_col = c;
  }
  t->total_walk = (int64_t)L.walk.size();
  t->sub_F.assign(L.na, 0);
  t->sub_qin.assign(L.na, {});
  for (int a = 0; a < L.na; ++a)
    if (d->site_active[a]) {
      const int NB = L.nb_ptr[a + 1] - L.nb_ptr[a], nspec = L.feat_ptr[a + 1] - L.feat_ptr[a];
      t->sub_F[a] = 2 * NB + nspec;
      // |e.e_n|, (e.e_n)^2, triple products and spin components are at most 1,
      // the three-pairing sums of four-body specs at most 3
      std::vector<float>& qs = t->sub_qin[a];
      qs.assign((size_t)t->sub_F[a], SUR_QIN_UNIT);
      for (int k = 0; k < nspec; ++k)
        if (L.feat[4 * (size_t)(L.feat_ptr[a] + k)] == 2) qs[(size_t)2 * NB + k] = SUR_QIN_PAIRS;
    }
  t->w_max = L.w_max;
  *out = t;
  return SCE_OK;
}

extern "C" int32_t sce_layout_check(const sce_tables_desc* d, int64_t o[5]) {
  if (!d || !o) return sce_fail(SCE_EINVAL, "sce_layout_check: NULL argument");
  HostLayout L;
  CHK(sce_derive_layout(d, &L));
  o[0] = L.na;
  o[1] = L.nb_max;
  o[2] = L.w_max;
  o[3] = (int64_t)L.walk.size();
  o[4] = L.nfeat_max;
  return SCE_OK;
}

extern "C" int32_t sce_tables_destroy(sce_tables* t) {
  delete t;
  return SCE_OK;
}

extern "C" int32_t sce_tables_position_tables(sce_tables* t, int32_t* nbq, uint64_t* low, uint16_t* perm,
                                              int64_t sizes[3]) {
  if (!t || !sizes) return sce_fail(SCE_EINVAL, "NULL argument");
  const DevTables& D = t->dev;
  sizes[0] = D.nb_max;
  sizes[1] = D.adj_max;
  sizes[2] = D.nbq_low ? 1 : 0;
  const size_t n = (size_t)D.n_active;
  if (nbq && n && D.nb_max) HIPCHK(hipMemcpy(nbq, D.nbq, n * D.nb_max * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (low && n && D.nbq_low) HIPCHK(hipMemcpy(low, D.nbq_low, n * sizeof(uint64_t), hipMemcpyDeviceToHost));
  if (perm && n) HIPCHK(hipMemcpy(perm, D.adj_perm, n * D.adj_max * sizeof(uint16_t), hipMemcpyDeviceToHost));
  return SCE_OK;
}

extern "C" int32_t sce_position_tables_host(const sce_tables_desc* d, int32_t* nbq, uint64_t* low, uint16_t* perm,
                                            int64_t sizes[3]) {
  if (!d || !sizes) return sce_fail(SCE_EINVAL, "NULL argument");
  HostLayout L;
  CHK(sce_derive_layout(d, &L, true));
  std::vector<int32_t> nq;
  std::vector<unsigned long long> lw;
  std::vector<uint16_t> pm;
  int amax = 0;
  CHK(sce_position_tables_host(d, L, nq, lw, pm, &amax));
  sizes[0] = L.nb_max;
  sizes[1] = amax;
  sizes[2] = lw.empty() ? 0 : 1;
  if (nbq) std::copy(nq.begin(), nq.end(), nbq);
  if (low && !lw.empty()) std::copy(lw.begin(), lw.begin() + std::max(d->n_active, 0), low);
  if (perm) std::copy(pm.begin(), pm.end(), perm);
  return SCE_OK;
}

extern "C" int32_t sce_tables_layout_info(const sce_tables* t, int64_t o[5]) {
  if (!t || !o) return sce_fail(SCE_EINVAL, "NULL argument");
  o[0] = t->dev.na;
  o[1] = t->dev.nb_max;
  o[2] = t->w_max;
  o[3] = t->total_walk;
  o[4] = t->dev.nfeat_max;
  return SCE_OK;
}

// ---- launch helpers --------------------------------------------------------
static size_t walk_lds_bytes(const DevTables& D, int L) {
  return (size_t)(TPB / L) * (size_t)(D.nb_max * D.nlm + 1 + D.nlm + (L > 64 ? L - 64 : 0)) * sizeof(double);
}

// Lanes per site: 16, 32 or 64 everywhere; 128 and 256 (two and four
// wavefronts per site) for plain Metropolis sweeps with per-colour launches
// and the walk test hook.
static int32_t check_lanes(int L, bool allow_wide = false) {
  if ((L == 128 || L == 256) && allow_wide) return SCE_OK;
  if (L != 16 && L != 32 && L != 64)
    return sce_fail(SCE_EINVAL, "lanes must be 16, 32 or 64%s (got %d)",
                    allow_wide ? ", 128 or 256" : " here (128, 256: plain Metropolis sweeps only)", L);
  return SCE_OK;
}

static int32_t check_lds(const DevTables& D, int L) {
  if (walk_lds_bytes(D, L) > 65536 || !D.walkp)
    return sce_fail(SCE_EINVAL, "neighbour stencil of %d sites needs %zu B of LDS per block at "
                    "lanes = %d; use more lanes", D.nb_max, walk_lds_bytes(D, L), L);
  return SCE_OK;
}

#define SCE_DISPATCH_LMAX_R(r, lmax, F, ...)             \
  switch (lmax) {                                        \
    case 0: r = F<0>(__VA_ARGS__); break;                \
    case 1: r = F<1>(__VA_ARGS__); break;                \
    case 2: r = F<2>(__VA_ARGS__); break;                \
    case 3: r = F<3>(__VA_ARGS__); break;                \
    case 4: r = F<4>(__VA_ARGS__); break;                \
    case 5: r = F<5>(__VA_ARGS__); break;                \
    default: r = F<6>(__VA_ARGS__); break;               \
  }

#define SCE_DISPATCH_LMAX(lmax, F, ...)                  \
  switch (lmax) {                                        \
    case 0: F<0>(__VA_ARGS__); break;                    \
    case 1: F<1>(__VA_ARGS__); break;                    \
    case 2: F<2>(__VA_ARGS__); break;                    \
    case 3: F<3>(__VA_ARGS__); break;                    \
    case 4: F<4>(__VA_ARGS__); break;                    \
    case 5: F<5>(__VA_ARGS__); break;                    \
    default: F<6>(__VA_ARGS__); break;                   \
  }

template <int LMAX>
static void launch_propose(const DevTables& D, const double* config, double* prop,
                           const ChainDev* st, hipStream_t s) {
  const int nb = (int)(((long long)D.n_active * PROPOSE_LANES + 255) / 256);
  if (nb) k_propose<LMAX><<<nb, 256, 0, s>>>(D, config, prop, st, 0, D.n_active, 1);
}


// With events e0/e1 the launch goes through hipExtLaunchKernelGGL, whose
// events carry the dispatch's own start and end timestamps (the kernel's
// execution time, without the queueing that separately recorded events add).
template <int L, int LMAX, bool TEST>
static void launch_metro_l(const DevTables& D, double* config, double* zrows, const double* prop,
                           double* dE, int* acc, const ChainDev* st, int q_lo, int q_n,
                           const int* ts, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  constexpr int G = TPB / L;
  const int nb = (q_n + G - 1) / G;
  if (!nb) return;
  if (e0)
    hipExtLaunchKernelGGL(k_metro<L, LMAX, TEST>, dim3(nb), dim3(TPB), (uint32_t)walk_lds_bytes(D, L), s, e0, e1,
                          0, D, config, zrows, prop, dE, acc, st, q_lo, q_n, ts, (double*)nullptr);
  else
    k_metro<L, LMAX, TEST><<<nb, TPB, walk_lds_bytes(D, L), s>>>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, ts);
}

template <int LMAX, bool TEST>
static void launch_metro(int L, const DevTables& D, double* config, double* zrows,
                         const double* prop, double* dE, int* acc, const ChainDev* st, int q_lo,
                         int q_n, const int* ts, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  if (L == 16) launch_metro_l<16, LMAX, TEST>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, ts, s, e0, e1);
  else if (L == 32) launch_metro_l<32, LMAX, TEST>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, ts, s, e0, e1);
  else if (L == 64) launch_metro_l<64, LMAX, TEST>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, ts, s, e0, e1);
  else if (L == 128) launch_metro_l<128, LMAX, TEST>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, ts, s, e0, e1);
  else launch_metro_l<256, LMAX, TEST>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, ts, s, e0, e1);
}
template <int LMAX> static void launch_metro_run(int L, const DevTables& D, double* config, double* zrows, const double* prop, double* dE, int* acc, const ChainDev* st, int q_lo, int q_n, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  launch_metro<LMAX, false>(L, D, config, zrows, prop, dE, acc, st, q_lo, q_n, nullptr, s, e0, e1);
}
template <int LMAX>
static void launch_metro_fin(int L, const DevTables& D, double* config, double* zrows, const double* prop, double* dE,
                             int* acc, ChainDev* st, int q_lo, int q_n, const FinalArgs& f, hipStream_t s) {
  const int nb = (q_n + TPB / L - 1) / (TPB / L) + 1;
  const size_t lb = walk_lds_bytes(D, L);
  if (L == 16) k_metro_fin<16, LMAX><<<nb, TPB, lb, s>>>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, f);
  else if (L == 32) k_metro_fin<32, LMAX><<<nb, TPB, lb, s>>>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, f);
  else if (L == 64) k_metro_fin<64, LMAX><<<nb, TPB, lb, s>>>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, f);
  else if (L == 128) k_metro_fin<128, LMAX><<<nb, TPB, lb, s>>>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, f);
  else k_metro_fin<256, LMAX><<<nb, TPB, lb, s>>>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, f);
}
template <int LMAX>
static void launch_metro_prop(int L, const DevTables& D, double* config, double* zrows, double* prop, double* dE,
                              int* acc, const ChainDev* st, int q_lo, int q_n, PropRange r0, PropRange r1,
                              hipStream_t s) {
  const int nwalk = (q_n + TPB / L - 1) / (TPB / L);
  const int nb = nwalk + prop_blocks(r0.n) + prop_blocks(r1.n);
  const size_t lb = walk_lds_bytes(D, L);
  if (L == 16) k_metro_prop<16, LMAX><<<nb, TPB, lb, s>>>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, nwalk, r0, r1);
  else if (L == 32) k_metro_prop<32, LMAX><<<nb, TPB, lb, s>>>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, nwalk, r0, r1);
  else if (L == 64) k_metro_prop<64, LMAX><<<nb, TPB, lb, s>>>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, nwalk, r0, r1);
  else if (L == 128) k_metro_prop<128, LMAX><<<nb, TPB, lb, s>>>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, nwalk, r0, r1);
  else k_metro_prop<256, LMAX><<<nb, TPB, lb, s>>>(D, config, zrows, prop, dE, acc, st, q_lo, q_n, nwalk, r0, r1);
}
template <int LMAX>
static void launch_propose_range(const DevTables& D, const double* config, double* prop, const ChainDev* st,
                                 PropRange r, hipStream_t s) {
  const int nb = prop_blocks(r.n);
  if (nb) k_propose<LMAX><<<nb, 256, 0, s>>>(D, config, prop, st, r.lo, r.n, r.off);
}
template <int LMAX> static void launch_metro_test(int L, const DevTables& D, double* zrows, const double* prop, double* dE, int q_n, const int* ts, hipStream_t s) {
  launch_metro<LMAX, true>(L, D, nullptr, zrows, prop, dE, nullptr, nullptr, 0, q_n, ts, s);
}

template <int L, int LMAX, bool TEST>
static void launch_or_l(const DevTables& D, double* config, double* zrows, double* dE, int* acc,
                        int* att, const ChainDev* st, int pass, int q_lo, int q_n, const int* ts,
                        double* cout, hipStream_t s) {
  constexpr int G = TPB / L, NLM = (LMAX + 1) * (LMAX + 1);
  const int nb = (q_n + G - 1) / G;
  if (!nb) return;
  // the LDS slots where the staged blocks leave room for them
  const size_t slots = walk_lds_bytes(D, L) + (size_t)G * L * NLM * sizeof(double);
  if (slots <= 65536 && D.walkp)
    k_or<L, LMAX, TEST, true><<<nb, TPB, slots, s>>>(D, config, zrows, dE, acc, att, st, pass, q_lo, q_n, ts, cout);
  else
    k_or<L, LMAX, TEST, false><<<nb, TPB, walk_lds_bytes(D, L), s>>>(D, config, zrows, dE, acc, att, st, pass, q_lo,
                                                                    q_n, ts, cout);
}
template <int LMAX, bool TEST>
static void launch_or(int L, const DevTables& D, double* config, double* zrows, double* dE,
                      int* acc, int* att, const ChainDev* st, int pass, int q_lo, int q_n,
                      const int* ts, double* cout, hipStream_t s) {
  if (L == 16) launch_or_l<16, LMAX, TEST>(D, config, zrows, dE, acc, att, st, pass, q_lo, q_n, ts, cout, s);
  else if (L == 32) launch_or_l<32, LMAX, TEST>(D, config, zrows, dE, acc, att, st, pass, q_lo, q_n, ts, cout, s);
  else launch_or_l<64, LMAX, TEST>(D, config, zrows, dE, acc, att, st, pass, q_lo, q_n, ts, cout, s);
}
template <int LMAX> static void launch_or_run(int L, const DevTables& D, double* config, double* zrows, double* dE, int* acc, int* att, const ChainDev* st, int pass, int q_lo, int q_n, hipStream_t s) {
  launch_or<LMAX, false>(L, D, config, zrows, dE, acc, att, st, pass, q_lo, q_n, nullptr, nullptr, s);
}
template <int LMAX> static void launch_or_test(int L, const DevTables& D, double* zrows, int q_n, const int* ts, double* cout, hipStream_t s) {
  launch_or<LMAX, true>(L, D, nullptr, zrows, nullptr, nullptr, nullptr, nullptr, 0, 0, q_n, ts, cout, s);
}

template <int LMAX>
static void launch_renorm(const DevTables& D, double* config, double* zrows, hipStream_t s) {
  const int nb = (D.n_sites + 255) / 256;
  if (nb) k_renorm<LMAX><<<nb, 256, 0, s>>>(D, config, zrows);
}

// Deterministic reduction of x[0..n) (+ the mean spin when meas) into the chain
// state: k_chunks over max(#x chunks, #site chunks) blocks, then k_final.
template <int MODE>
static void launch_reduce(const double* x, long long n, const int* iacc, const int* iatt,
                          double* part, long long* ipart, const double* config,
                          const uint8_t* active, long long n_sites, int meas, ChainDev* st,
                          double* out, long long n_att, double n_div, hipStream_t s,
                          DevBinner* bin = nullptr) {
  const long long nbx = std::max<long long>((n + TREE_B - 1) / TREE_B, 1);
  const long long nchm = meas ? (n_sites + TREE_B - 1) / TREE_B : 0;
  const long long grid = std::max(nbx, nchm);
  double* scratch = part + nbx + 3 * std::max<long long>(nchm, 1);
  k_chunks<MODE><<<(unsigned)grid, 256, 0, s>>>(x, n, iacc, iatt, part, ipart, config, active, n_sites, meas);
  k_final<MODE><<<1, 256, 0, s>>>(part, n, ipart, n_sites, meas, st, out, n_att, n_div, scratch,
                                   meas ? bin : nullptr);
}

// launch_reduce<MODE_METRO> with the next sweep's proposals in the first level.
template <int LMAX, int MODE = MODE_METRO>
static void launch_reduce_propose(sce_chain* c, int meas, hipStream_t s, bool defer_final, PropRange r) {
  const DevTables& D = c->t->dev;
  const long long n = D.n_active;
  const long long nbx = std::max<long long>((n + TREE_B - 1) / TREE_B, 1);
  const long long nchm = meas ? (D.n_sites + TREE_B - 1) / TREE_B : 0;
  const long long nchunk = std::max(nbx, nchm);
  const long long nprop = prop_blocks(r.n);
  double* scratch = c->partial + nbx + 3 * std::max<long long>(nchm, 1);
  k_chunks_propose<LMAX, MODE><<<(unsigned)(nchunk + nprop), 256, 0, s>>>(c->dE, n, c->acc,
                                                                         (MODE == MODE_DA || MODE == MODE_OR) ? c->att
                                                                                                             : nullptr,
                                                                         c->partial, c->ip64,
                                                                   c->config, D.active, D.n_sites, meas, nchunk, D,
                                                                   c->prop, c->st, r);
  if (!defer_final)
    k_final<MODE><<<1, 256, 0, s>>>(c->partial, n, c->ip64, D.n_sites, meas, c->st, c->obs, D.n_active,
                                    (double)D.n_active, scratch, meas ? c->bin : nullptr);
}
template <int LMAX>
static void launch_reduce_propose_da(sce_chain* c, int meas, hipStream_t s, bool defer_final, PropRange r) {
  launch_reduce_propose<LMAX, MODE_DA>(c, meas, s, defer_final, r);
}

// The operands of sweep-final reduction (k_final<MODE_METRO>) of a chain whose
// sweep measured (meas) or not.
static FinalArgs final_args(sce_chain* c, int meas) {
  const DevTables& D = c->t->dev;
  const long long n = D.n_active;
  const long long nbx = std::max<long long>((n + TREE_B - 1) / TREE_B, 1);
  const long long nchm = meas ? (D.n_sites + TREE_B - 1) / TREE_B : 0;
  FinalArgs f;
  f.part = c->partial;
  f.ipart = c->ip64;
  f.out = c->obs;
  f.scratch = c->partial + nbx + 3 * std::max<long long>(nchm, 1);
  f.bin = meas ? c->bin : nullptr;
  f.n = n;
  f.n_sites = D.n_sites;
  f.n_att = D.n_active;
  f.n_div = (double)D.n_active;
  f.meas = meas;
  return f;
}

static int32_t chain_check(sce_chain* c) {
  if (!c || !c->t) return sce_fail(SCE_EINVAL, "NULL chain");
  return SCE_OK;
}

static hipEvent_t next_event(sce_chain* c) {
  if (c->ev_used == c->ev.size()) {
    hipEvent_t e;
    (void)hipEventCreate(&e);
    c->ev.push_back(e);
  }
  return c->ev[c->ev_used++];
}

static int32_t set_beta(sce_chain* c, double beta) {
  if (!(std::isfinite(beta) && beta >= 0.0)) return sce_fail(SCE_EINVAL, "beta must be finite and >= 0");
  if (!(c->beta_dev == beta)) {
    k_set_params<<<1, 1, 0, c->stream>>>(c->st, beta, 0.0, 0);
    c->beta_dev = beta;
  }
  return SCE_OK;
}

// Enqueue one Metropolis sweep (no synchronisation).  (Proposing each colour
// on a forked graph branch, to overlap it with the earlier colours' walks,
// measured 40% slower per sweep: the cross-branch dependencies cost more than
// the proposal kernel.)
// need_prop: propose this sweep first (otherwise the previous sweep's
// reduction did); next_prop: propose the next sweep in this sweep's reduction.
// Where a chained sweep's proposals and deferred final go (enqueue_metro,
// enqueue_da): launch k (0..K-1) walks cols[k]; at[k] are the proposal ranges
// it carries as extra blocks (at[K]: the reduction launch; at most two per
// walk launch, one for the reduction); before/after: standalone proposal
// launches; kf: the walk launch whose block carries the pending final.
struct ChainPlan {
  std::vector<int> cols;
  int K = 0, kf = -1;
  std::vector<std::vector<PropRange>> at;
  std::vector<PropRange> before, after;
};

// color_ptr: the n_colors + 1 class offsets (0-based); fin_col: the class whose
// walk carries a pending final.
static int32_t plan_chain(int n_colors, const int32_t* color_ptr, int fin_col, bool next_prop, int pending,
                          bool carry_last, bool walk_props, ChainPlan* P) {
  ChainPlan& c = *P;
  for (int col = 0; col < n_colors; ++col)
    if (color_ptr[col + 1] > color_ptr[col]) c.cols.push_back(col);
  const int K = c.K = (int)c.cols.size();
  if (pending >= 0)
    for (int k = 0; k < K; ++k)
      if (c.cols[k] == fin_col) c.kf = k;
  const int kf = c.kf;
  // A proposal for this sweep (next = 0) or the next (next = 1) made at launch
  // j (-1: before the walks, K: the reduction) keys on the completed-sweep
  // count plus this offset; before the deferred final has run it lags by one.
  auto off_at = [&](int j, int next) { return 1 + next + ((kf >= 0 && j < kf) ? 1 : 0); };
  auto range_of = [&](int k) {
    const int lo = color_ptr[c.cols[k]];
    return PropRange{lo, color_ptr[c.cols[k] + 1] - lo, 0};
  };
  c.at.assign((size_t)K + 1, {});
  auto put = [&](int j, PropRange r, int next) {
    r.off = off_at(j, next);
    std::vector<PropRange>& v = c.at[j];
    if (!v.empty() && v.back().off == r.off && v.back().lo + v.back().n == r.lo)
      v.back().n += r.n;  // the classes are contiguous in colour order
    else
      v.push_back(r);
  };
  if (carry_last && K > 0) {
    // this sweep's proposals of its last class, made before that class's walk
    int j = 0;
    while (j < K - 1 && j == kf) ++j;
    if (j < K - 1) {
      put(j, range_of(K - 1), 0);
    } else {
      PropRange r = range_of(K - 1);
      r.off = off_at(-1, 0);
      c.before.push_back(r);
    }
  }
  if (next_prop) {
    for (int k = 0; k < K; ++k) {
      if (walk_props && k == K - 1) break;  // carried into the next sweep
      int j = walk_props ? k + 1 : K;
      while (j < K && j == kf) ++j;
      put(j, range_of(k), 1);
    }
  }
  for (int j = 0; j < K; ++j)  // a walk launch carries at most two ranges
    while (c.at[j].size() > 2) {
      PropRange r = c.at[j].back();
      c.at[j].pop_back();
      if (r.off != off_at(j, 1)) return sce_fail(SCE_EINVAL, "internal: proposal placement");
      r.off = off_at(K, 1);
      c.after.push_back(r);
    }
  while (c.at[K].size() > 1) {
    c.after.push_back(c.at[K].back());
    c.at[K].pop_back();
  }
  return SCE_OK;
}

// One Metropolis sweep.  Chaining within a call (next_prop, and in the
// previous sweep when !need_prop):
//  - proposals: a class's proposals for the next sweep depend only on its own
//    spins after its walk, so they ride as extra blocks in a later walk launch
//    of the same sweep (k_metro_prop); the last class's ride in the next
//    sweep's first walk launch (carry_last), or all go into the reduction
//    launch in timing mode, whose events time the walks alone;
//  - the final reduction (defer_final; pending >= 0 with the previous sweep's
//    measurement flag) rides in block 0 of the next sweep's smallest-class
//    walk (k_metro_fin), a launch that then carries no proposals, because its
//    final block advances the completed-sweep count the proposals key on.
static int32_t enqueue_metro(sce_chain* c, int L, int meas = 0, bool need_prop = true, bool next_prop = false,
                             bool defer_final = false, int pending = -1, bool carry_last = false) {
  CHK(check_lds(c->t->dev, L));
  if (c->sweeps_done >= 2147483647LL - 1)
    return sce_fail(SCE_ERANGE_COUNTER, "sweep_index exhausted the 32-bit RNG counter word");
  const DevTables& D = c->t->dev;
  const sce_tables* t = c->t;
  hipStream_t s = c->stream;
  ChainPlan P;
  CHK(plan_chain(t->n_colors, t->color_ptr.data(), t->fin_col, next_prop, pending, carry_last, !c->timing, &P));
  const int K = P.K, kf = P.kf;
  const std::vector<int>& cols = P.cols;
  const std::vector<std::vector<PropRange>>& at = P.at;
  const std::vector<PropRange>& before = P.before;
  const std::vector<PropRange>& after = P.after;
  for (const PropRange& r : before) SCE_DISPATCH_LMAX(D.lmax, launch_propose_range, D, c->config, c->prop, c->st, r, s);
  if (need_prop) SCE_DISPATCH_LMAX(D.lmax, launch_propose, D, c->config, c->prop, c->st, s);
  for (int k = 0; k < K; ++k) {
    const int lo = t->color_ptr[cols[k]], n = t->color_ptr[cols[k] + 1] - lo;
    if (k == kf) {
      const FinalArgs f = final_args(c, pending);
      SCE_DISPATCH_LMAX(D.lmax, launch_metro_fin, L, D, c->config, c->zrows, c->prop, c->dE, c->acc, c->st, lo, n,
                        f, s);
    } else if (!at[k].empty()) {
      const PropRange none{0, 0, 0};
      SCE_DISPATCH_LMAX(D.lmax, launch_metro_prop, L, D, c->config, c->zrows, c->prop, c->dE, c->acc, c->st, lo, n,
                        at[k][0], at[k].size() > 1 ? at[k][1] : none, s);
    } else {
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (c->timing) {
        e0 = next_event(c);
        e1 = next_event(c);
      }
      SCE_DISPATCH_LMAX(D.lmax, launch_metro_run, L, D, c->config, c->zrows, c->prop, c->dE, c->acc, c->st, lo, n,
                        s, e0, e1);
    }
  }
  if (pending >= 0 && kf < 0) {  // no colour launch to carry it
    const FinalArgs f = final_args(c, pending);
    k_final<MODE_METRO><<<1, 256, 0, s>>>(f.part, f.n, f.ipart, f.n_sites, f.meas, c->st, f.out, f.n_att, f.n_div,
                                          f.scratch, f.bin);
  }
  if (next_prop) {
    const PropRange r = at[K].empty() ? PropRange{0, 0, 0} : at[K][0];
    SCE_DISPATCH_LMAX(D.lmax, launch_reduce_propose, c, meas, s, defer_final, r);
  } else {
    launch_reduce<MODE_METRO>(c->dE, D.n_active, c->acc, nullptr, c->partial, c->ip64, c->config,
                              D.active, D.n_sites, meas, c->st, c->obs, D.n_active, (double)D.n_active, s, c->bin);
  }
  for (const PropRange& r : after) SCE_DISPATCH_LMAX(D.lmax, launch_propose_range, D, c->config, c->prop, c->st, r, s);
  c->sweeps_done += 1;
  return SCE_OK;
}

template <int LMAX>
static void launch_reduce_propose_or(sce_chain* c, int meas, hipStream_t s, bool defer_final, PropRange r) {
  launch_reduce_propose<LMAX, MODE_OR>(c, meas, s, defer_final, r);
}

// One overrelaxation pass (every colour, then the reduction).  meas: the
// compound sweep's measurement rides on this pass's reduction (the reference
// measures after the whole compound sweep, run.jl:143-150, 191-200);
// next_prop: the next Metropolis sweep's proposals ride in its first level
// (every spin is final for the sweep, and the Metropolis final of this sweep
// has run, so they key on the completed count + 1).
static int32_t enqueue_or(sce_chain* c, int L, int pass, int meas = 0, bool next_prop = false) {
  const DevTables& D = c->t->dev;
  CHK(check_lds(D, L));
  hipStream_t s = c->stream;
  for (int col = 0; col < c->t->n_colors; ++col) {
    const int lo = c->t->color_ptr[col], n = c->t->color_ptr[col + 1] - lo;
    if (n == 0) continue;
    SCE_DISPATCH_LMAX(D.lmax, launch_or_run, L, D, c->config, c->zrows, c->dE, c->acc, c->att, c->st, pass, lo, n, s);
  }
  if (next_prop) {
    SCE_DISPATCH_LMAX(D.lmax, launch_reduce_propose_or, c, meas, s, false, PropRange{0, D.n_active, 1});
  } else {
    launch_reduce<MODE_OR>(c->dE, D.n_active, c->acc, c->att, c->partial, c->ip64, c->config, D.active,
                           D.n_sites, meas, c->st, meas ? c->obs : nullptr, D.n_active, (double)D.n_active, s,
                           meas ? c->bin : nullptr);
  }
  return SCE_OK;
}

// One compound sweep (run.jl:143-150): a Metropolis sweep (its proposals made
// unless the previous compound sweep's last reduction made them), then
// `npass` overrelaxation passes; meas and next_prop ride on the last pass.
static int32_t enqueue_compound(sce_chain* c, int L, int npass, int meas, bool need_prop, bool next_prop) {
  CHK(enqueue_metro(c, L, 0, need_prop, false, false, -1, false));
  for (int p = 0; p < npass; ++p) {
    const bool last = p + 1 == npass;
    CHK(enqueue_or(c, L, p, last ? meas : 0, last && next_prop));
  }
  return SCE_OK;
}

static void enqueue_energy(sce_chain* c, int mode) {
  const DevTables& D = c->t->dev;
  const int nb = (D.n_inst + 255) / 256;
  if (nb) k_inst_energy<<<nb, 256, 0, c->stream>>>(D, c->zrows, c->inst_e);
  if (mode == MODE_RENORM)
    launch_reduce<MODE_RENORM>(c->inst_e, D.n_inst, nullptr, nullptr, c->partial, c->ip64, c->config,
                               D.active, D.n_sites, 0, c->st, nullptr, 0, 1.0, c->stream);
  else
    launch_reduce<MODE_TOTAL>(c->inst_e, D.n_inst, nullptr, nullptr, c->partial, c->ip64, c->config,
                              D.active, D.n_sites, 0, c->st, nullptr, 0, 1.0, c->stream);
}

static void enqueue_renorm(sce_chain* c) {
  const DevTables& D = c->t->dev;
  SCE_DISPATCH_LMAX(D.lmax, launch_renorm, D, c->config, c->zrows, c->stream);
  enqueue_energy(c, MODE_RENORM);
}

static void enqueue_measure(sce_chain* c) {
  const DevTables& D = c->t->dev;
  launch_reduce<MODE_MEASURE>(c->dE, 0, nullptr, nullptr, c->partial, c->ip64, c->config, D.active,
                              D.n_sites, 1, c->st, c->obs, 0, (double)D.n_active, c->stream, c->bin);
}

// Capture one Metropolis sweep (plain and with the fused measurement) as
// hipGraphs.  Every kernel reads the sweep counter, β and step from the
// device-resident chain state, so a captured sweep replays unchanged.
static int32_t ensure_graph(sce_chain* c, int L, int k) {
  if (c->g_lanes != L) {
    c->drop_graphs();
    c->g_lanes = L;
  }
  if (c->gexec[k]) return SCE_OK;
  const long long before = c->sweeps_done;
  const bool chained = (k & 2) == 0;  // the previous sweep made proposals and deferred its final
  const int32_t r = capture_graph(c->stream, c->gexec[k], [&] {
    return enqueue_metro(c, L, k & 1, !chained, (k & 4) != 0, (k & 4) != 0, chained ? (k >> 3) & 1 : -1, chained);
  });
  c->sweeps_done = before;
  return r;
}

// Capture one compound sweep as a hipGraph (variant k of cgexec).
static int32_t ensure_compound_graph(sce_chain* c, int L, int npass, int k) {
  if (c->g_lanes != L || c->cg_or != npass) {
    c->drop_graphs();
    c->g_lanes = L;
    c->cg_or = npass;
  }
  if (c->cgexec[k]) return SCE_OK;
  const long long before = c->sweeps_done;
  const int32_t r = capture_graph(c->stream, c->cgexec[k],
                                  [&] { return enqueue_compound(c, L, npass, k & 1, (k & 2) != 0, (k & 4) != 0); });
  c->sweeps_done = before;
  return r;
}

static int32_t read_state(sce_chain* c, ChainDev* h) {
  HIPCHK(hipMemcpyAsync(h, c->st, sizeof(ChainDev), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SCE_OK;
}

extern "C" int32_t sce_chain_create(sce_tables* t, const double* config, const double* zrows,
                                    double energy, uint64_t seed, double step, sce_chain** out) {
  if (!t || !config || !zrows || !out) return sce_fail(SCE_EINVAL, "sce_chain_create: NULL argument");
  *out = nullptr;
  if (!(step > 0.0)) return sce_fail(SCE_EINVAL, "step must be > 0; got %g", step);
  const DevTables& D = t->dev;
  auto* c = new sce_chain();
  c->t = t;
  auto fail = [&](int32_t r) {
    delete c;
    return r;
  };
  int32_t r;
  const size_t N = (size_t)D.n_sites, nlm = (size_t)D.nlm, na = (size_t)std::max(D.n_active, 1);
  if ((r = dupload(c->mem, &c->config, config, 3 * N)) != SCE_OK) return fail(r);
  if ((r = dupload(c->mem, &c->zrows, zrows, nlm * N)) != SCE_OK) return fail(r);
  if ((r = dalloc(c->mem, &c->prop, na * (nlm + 4))) != SCE_OK) return fail(r);
  if ((r = dalloc(c->mem, &c->dE, na)) != SCE_OK) return fail(r);
  if ((r = dalloc(c->mem, &c->acc, na)) != SCE_OK) return fail(r);
  if ((r = dalloc(c->mem, &c->att, na)) != SCE_OK) return fail(r);
  if ((r = dalloc(c->mem, &c->inst_e, (size_t)std::max(D.n_inst, 1))) != SCE_OK) return fail(r);
  const long long nmax = std::max<long long>({(long long)D.n_active, (long long)D.n_inst, (long long)D.n_sites});
  if ((r = dalloc(c->mem, &c->partial, partial_capacity(nmax) + 4 * ((size_t)nmax / TREE_B + 4))) != SCE_OK)
    return fail(r);
  if ((r = dalloc(c->mem, &c->ipart, (size_t)nmax / TREE_B + 4)) != SCE_OK) return fail(r);
  if ((r = dalloc(c->mem, &c->ip64, 2 * ((size_t)nmax / TREE_B + 4))) != SCE_OK) return fail(r);
  if ((r = dalloc(c->mem, &c->err, 4)) != SCE_OK) return fail(r);
  if (hipMemset(c->err, 0, 4 * sizeof(int)) != hipSuccess)
    return fail(sce_fail(SCE_EHIP, "hipMemset of the error flag failed"));
  if ((r = dalloc(c->mem, &c->st, 1)) != SCE_OK) return fail(r);
  ChainDev h{};
  h.energy = energy;
  h.step = step;
  h.beta = 0.0;
  h.seed = seed;
  if (hipMemcpy(c->st, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess)
    return fail(sce_fail(SCE_EHIP, "upload of chain state failed"));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(sce_fail(SCE_EHIP, "hipStreamCreate failed"));
  c->beta_dev = 0.0;
  *out = c;
  return SCE_OK;
}

extern "C" int32_t sce_chain_destroy(sce_chain* c) {
  if (c && c->stream) (void)hipStreamSynchronize(c->stream);
  delete c;
  return SCE_OK;
}

extern "C" int32_t sce_chain_info_get(sce_chain* c, sce_chain_info* o) {
  CHK(chain_check(c));
  if (!o) return sce_fail(SCE_EINVAL, "NULL info");
  ChainDev h;
  CHK(read_state(c, &h));
  o->energy = h.energy;
  o->step = h.step;
  o->max_drift = h.max_drift;
  o->last_drift = h.last_drift;
  o->sweep_index = h.sweep_index;
  o->acc_metro = h.acc_metro;
  o->att_metro = h.att_metro;
  o->acc_or = h.acc_or;
  o->att_or = h.att_or;
  o->seed = h.seed;
  o->acc_stage1 = h.acc_stage1;
  return SCE_OK;
}

extern "C" int32_t sce_chain_set_step(sce_chain* c, double step) {
  CHK(chain_check(c));
  if (!(step > 0.0)) return sce_fail(SCE_EINVAL, "step must be > 0");
  const double b = std::isnan(c->beta_dev) ? 0.0 : c->beta_dev;
  k_set_params<<<1, 1, 0, c->stream>>>(c->st, b, step, 1);
  HIPCHK(hipGetLastError());
  return SCE_OK;
}

extern "C" int32_t sce_chain_reset_counters(sce_chain* c) {
  CHK(chain_check(c));
  ChainDev h;
  CHK(read_state(c, &h));
  h.acc_metro = h.att_metro = h.acc_or = h.att_or = h.acc_stage1 = 0;
  h.max_drift = 0.0;
  HIPCHK(hipMemcpy(c->st, &h, sizeof h, hipMemcpyHostToDevice));
  return SCE_OK;
}

extern "C" int32_t sce_chain_set_counters(sce_chain* c, int64_t sweep_index, int64_t acc_metro,
                                          int64_t att_metro, int64_t acc_or, int64_t att_or) {
  CHK(chain_check(c));
  if (sweep_index < 0 || sweep_index >= 2147483647LL - 1)
    return sce_fail(SCE_ERANGE_COUNTER, "sweep_index %lld outside the 32-bit sweep counter", (long long)sweep_index);
  if (acc_metro < 0 || att_metro < acc_metro || acc_or < 0 || att_or < acc_or)
    return sce_fail(SCE_EINVAL, "counters must satisfy 0 <= acc <= att");
  ChainDev h;
  CHK(read_state(c, &h));
  h.sweep_index = sweep_index;
  h.acc_metro = acc_metro;
  h.att_metro = att_metro;
  h.acc_or = acc_or;
  h.att_or = att_or;
  HIPCHK(hipMemcpy(c->st, &h, sizeof h, hipMemcpyHostToDevice));
  c->sweeps_done = sweep_index;
  return SCE_OK;
}

extern "C" int32_t sce_chain_set_drift(sce_chain* c, double max_drift) {
  CHK(chain_check(c));
  if (!(max_drift >= 0.0)) return sce_fail(SCE_EINVAL, "max_drift must be >= 0");
  ChainDev h;
  CHK(read_state(c, &h));
  h.max_drift = max_drift;
  HIPCHK(hipMemcpy(c->st, &h, sizeof h, hipMemcpyHostToDevice));
  return SCE_OK;
}

extern "C" int32_t sce_download(sce_chain* c, double* config, double* zrows, double* energy) {
  CHK(chain_check(c));
  const DevTables& D = c->t->dev;
  if (config)
    HIPCHK(hipMemcpyAsync(config, c->config, 3 * (size_t)D.n_sites * 8, hipMemcpyDeviceToHost, c->stream));
  if (zrows)
    HIPCHK(hipMemcpyAsync(zrows, c->zrows, (size_t)D.nlm * D.n_sites * 8, hipMemcpyDeviceToHost, c->stream));
  ChainDev h;
  CHK(read_state(c, &h));
  if (energy) *energy = h.energy;
  return SCE_OK;
}

extern "C" int32_t sce_upload(sce_chain* c, const double* config, const double* zrows, double energy) {
  CHK(chain_check(c));
  const DevTables& D = c->t->dev;
  ChainDev h;
  CHK(read_state(c, &h));
  if (config) HIPCHK(hipMemcpy(c->config, config, 3 * (size_t)D.n_sites * 8, hipMemcpyHostToDevice));
  if (zrows) HIPCHK(hipMemcpy(c->zrows, zrows, (size_t)D.nlm * D.n_sites * 8, hipMemcpyHostToDevice));
  h.energy = energy;
  HIPCHK(hipMemcpy(c->st, &h, sizeof h, hipMemcpyHostToDevice));
  return SCE_OK;
}

extern "C" int32_t sce_swap_payload(sce_chain* a, sce_chain* b) {
  CHK(chain_check(a));
  CHK(chain_check(b));
  if (a->t != b->t) return sce_fail(SCE_EINVAL, "swap_payload: chains on different tables");
  HIPCHK(hipStreamSynchronize(a->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  ChainDev ha, hb;
  CHK(read_state(a, &ha));
  CHK(read_state(b, &hb));
  // ownership moves with the buffers: each chain frees what it now sweeps on
  DevBuf::exchange(a->mem, a->config, b->mem, b->config);
  DevBuf::exchange(a->mem, a->zrows, b->mem, b->zrows);
  std::swap(a->config, b->config);
  std::swap(a->zrows, b->zrows);
  a->drop_graphs();
  b->drop_graphs();
  std::swap(ha.energy, hb.energy);
  HIPCHK(hipMemcpy(a->st, &ha, sizeof ha, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->st, &hb, sizeof hb, hipMemcpyHostToDevice));
  return SCE_OK;
}

extern "C" int32_t sce_metropolis_sweep(sce_chain* c, double beta, int32_t lanes, int64_t* nacc) {
  CHK(chain_check(c));
  CHK(check_lanes(lanes, !c->persistent));
  CHK(set_beta(c, beta));
  c->timing = false;
  CHK(enqueue_metro(c, lanes));
  HIPCHK(hipGetLastError());
  ChainDev h;
  CHK(read_state(c, &h));
  if (nacc) *nacc = h.nacc_last;
  return SCE_OK;
}

extern "C" int32_t sce_overrelaxation_sweep(sce_chain* c, double beta, int32_t or_pass,
                                            int32_t lanes, int64_t* nacc) {
  CHK(chain_check(c));
  CHK(check_lanes(lanes));
  if (or_pass < 0 || or_pass > 1000) return sce_fail(SCE_EINVAL, "or_pass out of range");
  CHK(set_beta(c, beta));
  CHK(enqueue_or(c, lanes, or_pass));
  HIPCHK(hipGetLastError());
  ChainDev h;
  CHK(read_state(c, &h));
  if (nacc) *nacc = h.nacc_last;
  return SCE_OK;
}

// Segment buffers of the dependency-driven sweep for `cap` sweeps of which at
// most `mcap` are measured.
static constexpr int FLOW_SEG = 128;
static int32_t flow_reserve(sce_chain* c, int cap, int mcap) {
  sce_chain::Flow& F = c->flow;
  const DevTables& D = c->t->dev;
  const size_t n = (size_t)std::max(D.n_active, 1), N = (size_t)std::max(D.n_sites, 1);
  const size_t nbx = std::max<size_t>((n + TREE_B - 1) / TREE_B, 1), nchm = (N + TREE_B - 1) / TREE_B;
  if (!F.done) {
    CHK(dalloc(c->mem, &F.done, N));
    CHK(dalloc(c->mem, &F.scratch, partial_capacity((long long)std::max(n, N))));
  }
  if (cap > F.cap) {
    c->mem.release(F.dE);
    c->mem.release(F.acc);
    c->mem.release(F.part);
    c->mem.release(F.ipart);
    F.dE = nullptr, F.acc = nullptr, F.part = nullptr, F.ipart = nullptr;
    F.cap = 0;
    CHK(dalloc(c->mem, &F.dE, (size_t)cap * n));
    CHK(dalloc(c->mem, &F.acc, (size_t)cap * n));
    CHK(dalloc(c->mem, &F.part, (size_t)cap * (nbx + 3 * nchm)));
    CHK(dalloc(c->mem, &F.ipart, (size_t)cap * 2 * nbx));
    F.cap = cap;
  }
  if (mcap > F.mcap) {
    c->mem.release(F.cfg);
    F.cfg = nullptr;
    F.mcap = 0;
    CHK(dalloc(c->mem, &F.cfg, (size_t)mcap * 3 * N));
    F.mcap = mcap;
  }
  return SCE_OK;
}

// Workgroups of one wavefront of kernel f with `shmem` bytes of LDS that are
// resident on a CU at once.  The occupancy query (and the cooperative launch,
// which checks against it) overstates LDS-limited residency on this MI355X:
// a probe in which every wave waits, bounded, for all others found e.g. 9
// resident per CU at 15.5-16 KB where the query says 10, 25 at 5.5-6 KB
// (query 26-29) and 4 at 31.5-32 KB (query 5); every size from 1 to 64 KB
// holds at least 136 KB / shmem.  The dependency-driven sweeps need their
// whole grid resident, so they size it by the smaller bound.
static int32_t resident_per_cu(const void* f, size_t shmem, int* per_cu, int* cus) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, f, 64, shmem));
  constexpr size_t LDS_RESIDENT = 136 * 1024;
  if (shmem > 0) *per_cu = std::min<int>(*per_cu, (int)(LDS_RESIDENT / shmem));
  return SCE_OK;
}

template <int LMAX>
static int32_t launch_flow(sce_chain* c, FlowArgs& fa, hipStream_t s) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1);
  const DevTables& D = c->t->dev;
  const size_t shmem = ((size_t)D.nb_max * NLM + 1 + NLM) * sizeof(double);
  const void* f = reinterpret_cast<const void*>(&k_flow<LMAX>);
  int per_cu = 0, cus = 0;
  CHK(resident_per_cu(f, shmem, &per_cu, &cus));
  const long long grid = std::min<long long>((long long)per_cu * cus, std::max(D.n_active, 1));
  if (grid < 1) return sce_fail(SCE_EHIP, "the persistent sweep kernel cannot be resident");
  void* args[] = {&fa};
  HIPCHK(hipLaunchCooperativeKernel(f, dim3((unsigned)grid), dim3(64), args, (unsigned)shmem, s));
  return SCE_OK;
}

// n sweeps (sweeps i0 + 1 .. i0 + n of this call) as one dependency-driven
// launch plus the segment's reductions.
static int32_t enqueue_persistent(sce_chain* c, int L, double beta, long long n, long long i0,
                                  int64_t measure_interval, int defer_last) {
  const DevTables& D = c->t->dev;
  if (L != 64) return sce_fail(SCE_EINVAL, "persistent sweeps walk with 64 lanes per site; got %d", L);
  if (!D.nbq_low) return sce_fail(SCE_EINVAL, "persistent sweeps need at most 64 neighbour slots; the stencil has %d",
                                  D.nb_max);
  if (n <= 0) return SCE_OK;
  if (c->sweeps_done + n >= 2147483647LL - 1)
    return sce_fail(SCE_ERANGE_COUNTER, "sweep_index would exhaust the 32-bit RNG counter word");
  const int mi = (int)std::max<int64_t>(measure_interval, 0);
  int nmeas = 0;
  for (int i = 1; i <= n; ++i) nmeas += flow_meas_slot(i0, i, (int)n, mi, defer_last) >= 0 ? 1 : 0;
  CHK(flow_reserve(c, (int)n, nmeas));
  ChainDev h;
  CHK(read_state(c, &h));
  const sce_chain::Flow& F = c->flow;
  hipStream_t s = c->stream;
  const int na = D.n_active, N = D.n_sites;
  // first sweep's proposals; every site's done flag = the completed sweeps
  // (inactive sites never update: always done)
  SCE_DISPATCH_LMAX(D.lmax, launch_propose, D, c->config, c->prop, c->st, s);
  k_flow_init<<<(N + 255) / 256, 256, 0, s>>>(F.done, N, (int)c->sweeps_done);
  if (D.n_active < D.n_sites) k_flow_inactive<<<(N + 255) / 256, 256, 0, s>>>(F.done, D.active, N);
  HIPCHK(hipMemsetAsync(c->err, 0, sizeof(int), s));
  FlowArgs fa{};
  fa.T = D;
  fa.config = c->config;
  fa.zrows = c->zrows;
  fa.prop = c->prop;
  fa.dE_seg = F.dE;
  fa.acc_seg = F.acc;
  fa.cfg_seg = F.cfg;
  fa.done = F.done;
  fa.low = D.nbq_low;
  fa.err = c->err;
  fa.t0 = c->sweeps_done;
  fa.nsweeps = (int)n;
  fa.i0 = i0;
  fa.measure_interval = mi;
  fa.defer_last = defer_last;
  fa.beta = beta;
  fa.step = h.step;
  fa.seed = h.seed;
  fa.spin_limit = FLOW_SPIN_LIMIT;
  int32_t r = SCE_OK;
  SCE_DISPATCH_LMAX_R(r, D.lmax, launch_flow, c, fa, s);
  CHK(r);
  // the reductions of every sweep of the segment (the per-colour path's trees)
  const long long nbx = std::max<long long>((na + TREE_B - 1) / TREE_B, 1), nchm = (N + TREE_B - 1) / TREE_B;
  const long long per = nmeas > 0 ? std::max(nbx, nchm) : nbx;
  const long long pstride = nbx + 3 * nchm, istride = 2 * nbx;
  k_flow_chunks<<<(unsigned)(per * n), 256, 0, s>>>(F.dE, F.acc, F.cfg, D.active, na, N, F.part, F.ipart, pstride,
                                                     istride, (int)per, (int)n, i0, mi, defer_last);
  k_flow_final<<<1, 256, 0, s>>>(F.part, F.ipart, pstride, istride, na, N, c->st, c->obs, F.scratch, c->bin, (int)n,
                                 i0, mi, defer_last);
  HIPCHK(hipGetLastError());
  c->sweeps_done += n;
  return SCE_OK;
}

extern "C" int32_t sce_chain_set_persistent(sce_chain* c, int32_t on) {
  CHK(chain_check(c));
  c->persistent = on != 0;
  return SCE_OK;
}

extern "C" int32_t sce_chain_binner_reset(sce_chain* c, int32_t bin_size, int32_t nbins) {
  CHK(chain_check(c));
  if (bin_size < 0 || nbins < 0 || nbins > BIN_MAXSTORE || (bin_size > 0) != (nbins > 0) || nbins == 1)
    return sce_fail(SCE_EINVAL, "binner: need bin_size >= 1 and 2 <= nbins <= %d, or both 0", BIN_MAXSTORE);
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!c->bin_mem) CHK(dalloc(c->mem, &c->bin_mem, 1));
  DevBinner* h = new DevBinner();
  std::memset(h, 0, sizeof *h);
  h->bin_size = bin_size;
  h->nbins = nbins;
  const hipError_t e = hipMemcpy(c->bin_mem, h, sizeof *h, hipMemcpyHostToDevice);
  delete h;
  if (e != hipSuccess) return sce_fail(SCE_EHIP, "binner upload failed");
  if (!c->bin) c->drop_graphs();  // captured sweeps carry the binner pointer
  c->bin = c->bin_mem;
  return SCE_OK;
}

extern "C" int32_t sce_chain_binner_off(sce_chain* c) {
  CHK(chain_check(c));
  if (c->bin) c->drop_graphs();
  c->bin = nullptr;
  return SCE_OK;
}

extern "C" int32_t sce_chain_binner_get(sce_chain* c, int64_t* n, int64_t* count, double* sums, double* sums2,
                                        int32_t* nfull, double* means, int32_t means_cap) {
  CHK(chain_check(c));
  if (!c->bin_mem) return sce_fail(SCE_EINVAL, "binner: never reset on this chain");
  if (!n || !count || !sums || !sums2 || !nfull) return sce_fail(SCE_EINVAL, "binner: NULL output");
  HIPCHK(hipStreamSynchronize(c->stream));
  DevBinner* h = new DevBinner();
  const hipError_t e = hipMemcpy(h, c->bin_mem, sizeof *h, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    delete h;
    return sce_fail(SCE_EHIP, "binner download failed");
  }
  if (h->nfull > 0 && (!means || means_cap < h->nfull)) {
    delete h;
    return sce_fail(SCE_EDIM, "binner: means buffer holds %d bins; %d needed", means_cap, h->nfull);
  }
  *n = h->n;
  *nfull = h->nfull;
  for (int k = 0; k < BIN_LEVELS; ++k) {
    count[k] = h->count[k];
    for (int j = 0; j < BIN_COMP; ++j) {
      sums[k * BIN_COMP + j] = h->sums[k][j];
      sums2[k * BIN_COMP + j] = h->sums2[k][j];
    }
  }
  for (int i = 0; i < h->nfull; ++i)
    for (int j = 0; j < BIN_COMP; ++j) means[i * BIN_COMP + j] = h->means[i][j];
  delete h;
  return SCE_OK;
}

extern "C" int32_t sce_chain_set_graph(sce_chain* c, int32_t on) {
  CHK(chain_check(c));
  c->use_graph = on != 0;
  return SCE_OK;
}

extern "C" int32_t sce_run_sweeps(sce_chain* c, double beta, int64_t nsweeps, int32_t or_per_metropolis,
                                  int64_t renorm_interval, int64_t measure_interval, int32_t lanes,
                                  double* obs, int64_t obs_cap, int64_t* n_obs) {
  CHK(chain_check(c));
  // 128 and 256 lanes: plain sweeps with per-colour launches (no
  // overrelaxation pass or persistent grid at those widths)
  CHK(check_lanes(lanes, or_per_metropolis == 0 && !c->persistent));
  if (nsweeps < 0) return sce_fail(SCE_EINVAL, "nsweeps must be >= 0");
  if (or_per_metropolis < 0) return sce_fail(SCE_EINVAL, "or_per_metropolis must be >= 0");
  const int64_t want_obs = measure_interval > 0 ? nsweeps / measure_interval : 0;
  if (want_obs > 0 && (!obs || obs_cap < want_obs))
    return sce_fail(SCE_EDIM, "obs buffer holds %lld rows; %lld needed", (long long)obs_cap, (long long)want_obs);
  if (c->sweeps_done + nsweeps >= 2147483647LL - 1)
    return sce_fail(SCE_ERANGE_COUNTER, "sweep_index would exhaust the 32-bit RNG counter word");
  if (want_obs > 0 && (size_t)want_obs > c->obs_cap) {
    // captured sweeps hold the row buffer's address, so a new buffer drops
    // them: grow geometrically, from a size most calls never exceed
    const size_t cap = obs_rows_for((size_t)want_obs, c->obs_cap);
    double* p;
    CHK(dalloc(c->mem, &p, 4 * cap));
    c->drop_graphs();
    c->mem.release(c->obs);
    c->obs = p;
    c->obs_cap = cap;
  }
  CHK(set_beta(c, beta));
  // reset the observation counter
  k_reset_nobs<<<1, 1, 0, c->stream>>>(c->st);
  c->timing = c->timing_enabled;
  c->ev_used = 0;
  hipEvent_t t0 = next_event(c);
  HIPCHK(hipEventRecord(t0, c->stream));
  if (c->persistent && or_per_metropolis == 0) {
    // segments end at renormalisation points (and hold at most FLOW_SEG
    // sweeps); each is one dependency-driven cooperative launch
    int64_t i = 0;
    while (i < nsweeps) {
      int64_t seg = std::min<int64_t>(nsweeps - i, FLOW_SEG);
      if (renorm_interval > 0) seg = std::min<int64_t>(seg, renorm_interval - (i % renorm_interval));
      hipEvent_t e0 = nullptr;
      if (c->timing) {
        e0 = next_event(c);
        HIPCHK(hipEventRecord(e0, c->stream));
      }
      const bool renorm_after = renorm_interval > 0 && (i + seg) % renorm_interval == 0;
      const bool meas_end = measure_interval > 0 && (i + seg) % measure_interval == 0;
      CHK(enqueue_persistent(c, lanes, beta, seg, i, measure_interval, renorm_after && meas_end));
      if (c->timing) HIPCHK(hipEventRecord(next_event(c), c->stream));
      i += seg;
      if (renorm_after) {
        enqueue_renorm(c);
        if (meas_end)  // the reference measures after renormalising (run.jl:191-200)
          enqueue_measure(c);
      }
    }
  } else if (or_per_metropolis > 0) {
    // compound sweeps: each is one captured graph (Metropolis, then the
    // passes); the last pass's reduction carries the measurement and, unless a
    // renormalisation or the call's end comes next, the next sweep's proposals
    bool have_prop = false;
    for (int64_t i = 1; i <= nsweeps; ++i) {
      const bool renorm = renorm_interval > 0 && i % renorm_interval == 0;
      const bool meas = measure_interval > 0 && i % measure_interval == 0;
      const bool fused = meas && !renorm;
      const bool next_prop = !renorm && i < nsweeps;
      if (c->use_graph && !c->timing) {
        const int k = (fused ? 1 : 0) + (have_prop ? 0 : 2) + (next_prop ? 4 : 0);
        CHK(ensure_compound_graph(c, lanes, or_per_metropolis, k));
        HIPCHK(hipGraphLaunch(c->cgexec[k].x, c->stream));
        c->sweeps_done += 1;
      } else {
        CHK(enqueue_compound(c, lanes, or_per_metropolis, fused ? 1 : 0, !have_prop, next_prop));
      }
      have_prop = next_prop;
      if (renorm) enqueue_renorm(c);
      if (meas && !fused) enqueue_measure(c);
    }
  } else {
    const bool graphs = c->use_graph && !c->timing;
    // Proposals depend only on each site's own spin at the sweep's start, so a
    // sweep's reduction may make the next sweep's (fused launch) when nothing
    // changes spins in between: not across renormalisations or overrelaxation,
    // and never across calls (the step or state may change between them).
    // With the proposals, the sweep's final reduction moves into the next
    // sweep's smallest-class launch (block 0 of k_metro_fin): nothing between
    // the two reads or writes the chain scalars it updates.  Not in timing
    // mode, whose per-launch events time the walks alone (enqueue_metro).
    bool have_prop = false, carry = false;
    int pending = -1;
    for (int64_t i = 1; i <= nsweeps; ++i) {
      const bool renorm = renorm_interval > 0 && i % renorm_interval == 0;
      const bool meas = measure_interval > 0 && i % measure_interval == 0;
      // a measurement rides on the sweep's own reduction unless something
      // (overrelaxation, renormalisation) comes between them (run.jl:191-200)
      const bool fused = meas && !renorm;
      const bool next_prop = !renorm && i < nsweeps;
      const bool defer = next_prop && !c->timing;
      if (graphs) {
        const int k = (fused ? 1 : 0) + (have_prop ? 0 : 2) + (next_prop ? 4 : 0) + (pending == 1 ? 8 : 0);
        CHK(ensure_graph(c, lanes, k));
        HIPCHK(hipGraphLaunch(c->gexec[k].x, c->stream));
        c->sweeps_done += 1;
      } else {
        CHK(enqueue_metro(c, lanes, fused ? 1 : 0, !have_prop, next_prop, defer, pending, carry));
      }
      have_prop = next_prop;
      pending = defer ? (fused ? 1 : 0) : -1;
      carry = next_prop && !c->timing;
      if (renorm) enqueue_renorm(c);
      if (meas && !fused) enqueue_measure(c);
    }
  }
  hipEvent_t t1 = next_event(c);
  HIPCHK(hipEventRecord(t1, c->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->timing = false;
  if (c->persistent && or_per_metropolis == 0 && nsweeps > 0) {
    int e = 0;
    HIPCHK(hipMemcpy(&e, c->err, sizeof e, hipMemcpyDeviceToHost));
    if (e) return sce_fail(SCE_EHIP, "persistent sweep: a wait on neighbour updates exceeded its bound");
  }
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, t0, t1));
  c->total_ms = ms;
  c->walk_ms = 0.0;
  c->walk_launches = 0;
  for (size_t k = 1; k + 1 < c->ev_used; k += 2) {
    float m = 0.f;
    HIPCHK(hipEventElapsedTime(&m, c->ev[k], c->ev[k + 1]));
    c->walk_ms += m;
    c->walk_launches += 1;
  }
  if (want_obs > 0) {
    HIPCHK(hipMemcpy(obs, c->obs, 4 * (size_t)want_obs * 8, hipMemcpyDeviceToHost));
  }
  if (n_obs) *n_obs = want_obs;
  return SCE_OK;
}

extern "C" int32_t sce_chain_set_timing(sce_chain* c, int32_t on) {
  CHK(chain_check(c));
  c->timing_enabled = on != 0;
  return SCE_OK;
}

extern "C" int32_t sce_chain_timing(sce_chain* c, double* total_ms, double* walk_ms_avg,
                                    int64_t* walk_launches) {
  CHK(chain_check(c));
  if (total_ms) *total_ms = c->total_ms;
  if (walk_ms_avg) *walk_ms_avg = c->walk_launches ? c->walk_ms / (double)c->walk_launches : 0.0;
  if (walk_launches) *walk_launches = c->walk_launches;
  return SCE_OK;
}

extern "C" int32_t sce_renormalize(sce_chain* c, double* drift) {
  CHK(chain_check(c));
  enqueue_renorm(c);
  HIPCHK(hipGetLastError());
  ChainDev h;
  CHK(read_state(c, &h));
  if (drift) *drift = h.last_drift;
  return SCE_OK;
}

extern "C" int32_t sce_chain_total_energy(sce_chain* c, double* E) {
  CHK(chain_check(c));
  enqueue_energy(c, MODE_TOTAL);
  HIPCHK(hipGetLastError());
  ChainDev h;
  CHK(read_state(c, &h));
  if (E) *E = h.e_full;
  return SCE_OK;
}

// ---- diagnostics -------------------------------------------------------------
template <class T>
struct Scratch {
  T* p = nullptr;
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
  int32_t alloc(size_t n) {
    hipError_t e = hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T));
    return e == hipSuccess ? SCE_OK : sce_fail(SCE_EHIP, "hipMalloc: %s", hipGetErrorString(e));
  }
  int32_t up(const T* src, size_t n) {
    CHK(alloc(n));
    if (n) HIPCHK(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    return SCE_OK;
  }
  int32_t down(T* dst, size_t n) {
    if (n) HIPCHK(hipMemcpy(dst, p, n * sizeof(T), hipMemcpyDeviceToHost));
    return SCE_OK;
  }
};

// ---- all-site energy gradient (gpu_energy_gradient!, gpu/gpu_gradient.jl) -----
template <int LMAX>
static void launch_rows_of(int n, const double* config, double* zrows, hipStream_t s) {
  if (n > 0) k_rows_of<LMAX><<<(n + 127) / 128, 128, 0, s>>>(n, config, zrows);
}

template <int LMAX>
// The walk-layout kernel where its LDS (neighbour rows, gradient rows and
// share partials of four sites) fits one block, else the adjacency-table one;
// the two are bit-identical.
static void launch_grad(const GradTables& T, const DevTables& D, const double* config, const double* zrows, double* G,
                        int lanes, hipStream_t s) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1);
  const size_t lb = (size_t)(TPB / 64) * (D.nb_max * NLM + 1 + 3 * NLM + 3 * lanes) * sizeof(double);
  if (lb > 65536) {
    const int nb = (T.n_sites + TPB / 64 - 1) / (TPB / 64);
    if (nb == 0) return;
    if (lanes == 64) k_grad<LMAX, 64><<<nb, TPB, 0, s>>>(T, config, zrows, G);
    else if (lanes == 128) k_grad<LMAX, 128><<<nb, TPB, 0, s>>>(T, config, zrows, G);
    else k_grad<LMAX, 256><<<nb, TPB, 0, s>>>(T, config, zrows, G);
    return;
  }
  (void)hipMemsetAsync(G, 0, 3 * (size_t)D.n_sites * sizeof(double), s);  // inactive sites
  const int nb = (D.n_active + TPB / 64 - 1) / (TPB / 64);
  if (nb == 0) return;
  if (lanes == 64) k_grad_walk<LMAX, 64><<<nb, TPB, lb, s>>>(D, config, zrows, G);
  else if (lanes == 128) k_grad_walk<LMAX, 128><<<nb, TPB, lb, s>>>(D, config, zrows, G);
  else k_grad_walk<LMAX, 256><<<nb, TPB, lb, s>>>(D, config, zrows, G);
}

static int32_t check_grad_lanes(int lanes) {
  if (lanes != 64 && lanes != 128 && lanes != 256)
    return sce_fail(SCE_EINVAL, "gradient lanes (workgroup size) must be 64, 128 or 256 (got %d)", lanes);
  return SCE_OK;
}

extern "C" int32_t sce_energy_gradient(sce_tables* t, const double* config, double* G, int32_t lanes) {
  if (!t || !config || !G) return sce_fail(SCE_EINVAL, "sce_energy_gradient: NULL argument");
  CHK(check_grad_lanes(lanes));
  const DevTables& D = t->dev;
  const size_t n = (size_t)D.n_sites;
  Scratch<double> cf, z, g;
  CHK(cf.up(config, 3 * n));
  CHK(z.alloc((size_t)D.nlm * n));
  CHK(g.alloc(3 * n));
  SCE_DISPATCH_LMAX(D.lmax, launch_rows_of, D.n_sites, cf.p, z.p, (hipStream_t)0);
  SCE_DISPATCH_LMAX(D.lmax, launch_grad, t->grad, D, cf.p, z.p, g.p, lanes, (hipStream_t)0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return g.down(G, 3 * n);
}

// ---------------------------------------------------------------------------
// The all-site gradient without allocation: the device copy of the
// configuration, its rebuilt rows and the gradient live in one file-scope
// device array (each GPU holds its own copy for the life of the process).
// Every call on a GPU shares that GPU's array, so calls that overlap on one GPU
// race; the array keeps the last call's contents.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_DOUBLES = (size_t)4 << 20;  // 32 MiB
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

static int32_t scratch_ptr(double** out) {
  // resolved once per device, with that device current: the array's name is
  // a host-side stand-in, not a device address
  constexpr int MAXDEV = 64;
  static std::mutex mu;
  static double* ptrs[MAXDEV] = {};
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAXDEV) return sce_fail(SCE_EINVAL, "device %d beyond the %d the library tracks", dev, MAXDEV);
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return sce_fail(SCE_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    ptrs[dev] = (double*)p;
  }
  *out = ptrs[dev];
  return SCE_OK;
}

extern "C" int32_t sce_energy_gradient_noalloc(sce_tables* t, const double* config, double* G, int32_t lanes) {
  if (!t || !config || !G) return sce_fail(SCE_EINVAL, "sce_energy_gradient_noalloc: NULL argument");
  CHK(check_grad_lanes(lanes));
  const DevTables& D = t->dev;
  const size_t n = (size_t)D.n_sites;
  const size_t need = (3 + (size_t)D.nlm + 3) * n;
  if (need > NOALLOC_DOUBLES)
    return sce_fail(SCE_EDIM,
                    "sce_energy_gradient_noalloc: %zu sites at nlm %d need %zu doubles of scratch; the static "
                    "array holds %zu (use sce_energy_gradient)",
                    n, D.nlm, need, NOALLOC_DOUBLES);
  double* base = nullptr;
  CHK(scratch_ptr(&base));
  double* cf = base;                 // 3 x n
  double* z = cf + 3 * n;            // nlm x n
  double* g = z + (size_t)D.nlm * n; // 3 x n (zeroed for inactive sites by launch_grad)
  if (n) HIPCHK(hipMemcpy(cf, config, 3 * n * sizeof(double), hipMemcpyHostToDevice));
  SCE_DISPATCH_LMAX(D.lmax, launch_rows_of, D.n_sites, cf, z, (hipStream_t)0);
  SCE_DISPATCH_LMAX(D.lmax, launch_grad, t->grad, D, cf, z, g, lanes, (hipStream_t)0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  if (n) HIPCHK(hipMemcpy(G, g, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
  return SCE_OK;
}

extern "C" int32_t sce_chain_energy_gradient(sce_chain* c, double* G, int32_t lanes) {
  CHK(chain_check(c));
  if (!G) return sce_fail(SCE_EINVAL, "sce_chain_energy_gradient: NULL output");
  CHK(check_grad_lanes(lanes));
  const DevTables& D = c->t->dev;
  Scratch<double> g;
  CHK(g.alloc(3 * (size_t)D.n_sites));
  SCE_DISPATCH_LMAX(D.lmax, launch_grad, c->t->grad, D, c->config, c->zrows, g.p, lanes, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return g.down(G, 3 * (size_t)D.n_sites);
}

// Average duration of the gradient kernel over `reps` back-to-back launches on
// the chain's rows (HIP events on the chain's stream; nothing is changed).
extern "C" int32_t sce_chain_time_gradient(sce_chain* c, int32_t lanes, int32_t reps, double* ms_avg) {
  CHK(chain_check(c));
  CHK(check_grad_lanes(lanes));
  if (reps < 1 || !ms_avg) return sce_fail(SCE_EINVAL, "time_gradient: reps >= 1 and an output needed");
  const DevTables& D = c->t->dev;
  Scratch<double> g;
  CHK(g.alloc(3 * (size_t)D.n_sites));
  SCE_DISPATCH_LMAX(D.lmax, launch_grad, c->t->grad, D, c->config, c->zrows, g.p, lanes, c->stream);
  hipEvent_t e0 = next_event(c), e1 = next_event(c);
  c->ev_used = 0;
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; ++r)
    SCE_DISPATCH_LMAX(D.lmax, launch_grad, c->t->grad, D, c->config, c->zrows, g.p, lanes, c->stream);
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *ms_avg = (double)ms / reps;
  return SCE_OK;
}

// Average duration of the colour-walk kernel: `reps` passes over every colour
// class launched back to back on the chain's stream between one pair of HIP
// events (the GPU runs them without gaps, so elapsed / launches is the kernel's
// own duration).  The walks re-use the sweep's proposal buffer; config and rows
// are saved before and restored after, and no counter or energy changes.
extern "C" int32_t sce_chain_time_walk(sce_chain* c, int32_t lanes, int32_t reps, double* ms_avg,
                                       int64_t* launches) {
  CHK(chain_check(c));
  CHK(check_lanes(lanes, true));
  CHK(check_lds(c->t->dev, lanes));
  if (reps < 1) return sce_fail(SCE_EINVAL, "reps must be >= 1");
  const DevTables& D = c->t->dev;
  const size_t ncfg = 3 * (size_t)D.n_sites, nz = (size_t)D.nlm * D.n_sites;
  Scratch<double> cfg, z;
  CHK(cfg.alloc(ncfg));
  CHK(z.alloc(nz));
  HIPCHK(hipMemcpyAsync(cfg.p, c->config, ncfg * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(z.p, c->zrows, nz * 8, hipMemcpyDeviceToDevice, c->stream));
  SCE_DISPATCH_LMAX(D.lmax, launch_propose, D, c->config, c->prop, c->st, c->stream);
  hipEvent_t e0 = next_event(c), e1 = next_event(c);
  c->ev_used = 0;
  int64_t n = 0;
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; ++r)
    for (int col = 0; col < c->t->n_colors; ++col) {
      const int lo = c->t->color_ptr[col], cnt = c->t->color_ptr[col + 1] - lo;
      if (cnt == 0) continue;
      SCE_DISPATCH_LMAX(D.lmax, launch_metro_run, lanes, D, c->config, c->zrows, c->prop, c->dE, c->acc, c->st,
                        lo, cnt, c->stream);
      ++n;
    }
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipMemcpyAsync(c->config, cfg.p, ncfg * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->zrows, z.p, nz * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  if (ms_avg) *ms_avg = n ? ms / (double)n : 0.0;
  if (launches) *launches = n;
  return SCE_OK;
}


static int32_t need_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return sce_fail(SCE_ENODEV, "no HIP device available");
  return SCE_OK;
}

extern "C" int32_t sce_test_philox(int64_t n, const uint64_t* seeds, const uint32_t* ctr4, uint32_t* out4) {
  CHK(need_device());
  Scratch<uint64_t> s;
  Scratch<uint32_t> c, o;
  CHK(s.up(seeds, n));
  CHK(c.up(ctr4, 4 * n));
  CHK(o.alloc(4 * n));
  if (n) k_test_philox<<<(unsigned)((n + 255) / 256), 256>>>(n, s.p, c.p, o.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return o.down(out4, 4 * n);
}

extern "C" int32_t sce_test_chain_plan(int32_t n_colors, const int32_t* color_ptr, int32_t next_prop,
                                       int32_t pending, int32_t carry_last, int32_t walk_props, int32_t* out,
                                       int32_t cap, int32_t* n_out, int32_t* kf) {
  if (n_colors < 0 || !color_ptr || !out || !n_out || !kf) return sce_fail(SCE_EINVAL, "NULL argument");
  for (int c = 0; c < n_colors; ++c)
    if (color_ptr[c + 1] < color_ptr[c]) return sce_fail(SCE_EINVAL, "class offsets must not decrease");
  int fin_col = -1;  // the smallest non-empty class, as sce_tables_create picks it
  for (int c = 0; c < n_colors; ++c) {
    const int n = color_ptr[c + 1] - color_ptr[c];
    if (n > 0 && (fin_col < 0 || n < color_ptr[fin_col + 1] - color_ptr[fin_col])) fin_col = c;
  }
  ChainPlan P;
  CHK(plan_chain(n_colors, color_ptr, fin_col, next_prop != 0, pending, carry_last != 0, walk_props != 0, &P));
  int m = 0;
  auto emit = [&](int launch, const PropRange& r) {
    if (m < cap) {
      out[4 * m] = launch;
      out[4 * m + 1] = r.lo;
      out[4 * m + 2] = r.n;
      out[4 * m + 3] = r.off;
    }
    ++m;
  };
  for (const PropRange& r : P.before) emit(-1, r);
  for (int j = 0; j <= P.K; ++j)
    for (const PropRange& r : P.at[j]) emit(j, r);
  for (const PropRange& r : P.after) emit(P.K + 1, r);
  *n_out = m;
  *kf = P.kf;
  return m <= cap ? SCE_OK : sce_fail(SCE_EDIM, "%d placements, room for %d", m, cap);
}

extern "C" int32_t sce_test_detmath(int64_t n, int32_t op, const double* x, double* out) {
  CHK(need_device());
  Scratch<double> xi, o;
  CHK(xi.up(x, n));
  CHK(o.alloc(n));
  if (n) k_test_detmath<<<(unsigned)((n + 255) / 256), 256>>>(n, op, xi.p, o.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return o.down(out, n);
}

template <int LMAX>
static void launch_test_zlm(long long n, const double* d, double* o) {
  if (n) k_test_zlm<LMAX><<<(unsigned)((n + 255) / 256), 256>>>(n, d, o);
}

extern "C" int32_t sce_test_zlm(int64_t n, int32_t lmax, const double* dirs3, double* out) {
  CHK(need_device());
  if (lmax < 0 || lmax > 6) return sce_fail(SCE_EINVAL, "lmax out of range");
  const int nlm = (lmax + 1) * (lmax + 1);
  Scratch<double> d, o;
  CHK(d.up(dirs3, 3 * n));
  CHK(o.alloc((size_t)nlm * n));
  SCE_DISPATCH_LMAX(lmax, launch_test_zlm, n, d.p, o.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return o.down(out, (size_t)nlm * n);
}

extern "C" int32_t sce_test_proposal(int64_t n, uint64_t seed, const int32_t* sites, const int32_t* sweeps,
                                     const double* e3, double step, double* e2_out, double* u_out) {
  CHK(need_device());
  Scratch<int32_t> si, sw;
  Scratch<double> e, e2, u;
  CHK(si.up(sites, n));
  CHK(sw.up(sweeps, n));
  CHK(e.up(e3, 3 * n));
  CHK(e2.alloc(3 * n));
  CHK(u.alloc(n));
  if (n) k_test_proposal<<<(unsigned)((n + 255) / 256), 256>>>(n, seed, si.p, sw.p, e.p, step, e2.p, u.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  CHK(e2.down(e2_out, 3 * n));
  return u.down(u_out, n);
}

static int32_t check_sites(const sce_tables* t, int64_t n, const int32_t* sites, std::vector<int32_t>& z0) {
  z0.resize(n);
  std::vector<uint8_t> act(t->dev.n_sites);
  HIPCHK(hipMemcpy(act.data(), t->dev.active, act.size(), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; ++i) {
    if (sites[i] < 1 || sites[i] > t->dev.n_sites) return sce_fail(SCE_EINVAL, "site %d out of range", sites[i]);
    if (!act[sites[i] - 1]) return sce_fail(SCE_EINVAL, "site %d is inactive", sites[i]);
    z0[i] = sites[i] - 1;
  }
  return SCE_OK;
}

extern "C" int32_t sce_test_walk_delta(sce_tables* t, const double* zrows, int64_t n, const int32_t* sites,
                                       const double* znew, int32_t lanes, double* out) {
  if (!t) return sce_fail(SCE_EINVAL, "NULL tables");
  CHK(check_lanes(lanes, true));
  const DevTables& D = t->dev;
  CHK(check_lds(D, lanes));
  std::vector<int32_t> s0;
  CHK(check_sites(t, n, sites, s0));
  const int nlm = D.nlm;
  std::vector<double> prop((size_t)n * (nlm + 4), 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < nlm; ++k) prop[(size_t)i * (nlm + 4) + 4 + k] = znew[(size_t)i * nlm + k];
  Scratch<double> z, p, o;
  Scratch<int32_t> si;
  CHK(z.up(zrows, (size_t)nlm * D.n_sites));
  CHK(p.up(prop.data(), prop.size()));
  CHK(o.alloc(n));
  CHK(si.up(s0.data(), n));
  SCE_DISPATCH_LMAX(D.lmax, launch_metro_test, lanes, D, z.p, p.p, o.p, (int)n, si.p, nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return o.down(out, n);
}

extern "C" int32_t sce_test_walk_coeffs(sce_tables* t, const double* zrows, int64_t n, const int32_t* sites,
                                        int32_t lanes, double* out) {
  if (!t) return sce_fail(SCE_EINVAL, "NULL tables");
  CHK(check_lanes(lanes));
  const DevTables& D = t->dev;
  CHK(check_lds(D, lanes));
  std::vector<int32_t> s0;
  CHK(check_sites(t, n, sites, s0));
  const int nlm = D.nlm;
  Scratch<double> z, o;
  Scratch<int32_t> si;
  CHK(z.up(zrows, (size_t)nlm * D.n_sites));
  CHK(o.alloc((size_t)n * nlm));
  CHK(si.up(s0.data(), n));
  SCE_DISPATCH_LMAX(D.lmax, launch_or_test, lanes, D, z.p, (int)n, si.p, o.p, nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return o.down(out, (size_t)n * nlm);
}

extern "C" int32_t sce_test_reduce(int64_t n, const double* x, double* out) {
  CHK(need_device());
  Scratch<double> xi, part, o;
  Scratch<ChainDev> st;
  CHK(xi.up(x, n));
  CHK(part.alloc(partial_capacity(n)));
  CHK(o.alloc(1));
  ChainDev h{};
  CHK(st.up(&h, 1));
  Scratch<long long> ip;
  CHK(ip.alloc(2 * ((size_t)n / TREE_B + 4)));
  launch_reduce<MODE_TEST>(xi.p, n, nullptr, nullptr, part.p, ip.p, nullptr, nullptr, 0, 0, st.p, o.p, 0,
                           1.0, nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return o.down(out, 1);
}

// ===========================================================================
// Batches: many chains — replicas of one model (parallel tempering) or
// independent chains of different models and sizes with the same lmax — in ONE
// launch per colour index.  Each chain keeps its own keyed stream (seed, site,
// sweep), its own device state and its own deterministic reductions, so every
// chain of a batch is bit-identical to the same chain run alone.
//
// Replica exchange (src/gpu/gpu_pt.jl:46-63, src/pt.jl:92-96) runs on the
// device: the payload (config, rows, energy) of adjacent rungs is swapped by
// exchanging the payload pointers in the chain descriptors; seed, sweep index,
// step and counters stay with the rung.  The exchange uniform is keyed:
// block (exchange seed, pair i, exchange count, slot 4, 0), words 1-2.
// ===========================================================================
struct BChain {
  DevTables T;
  double *config, *zrows, *prop, *dE, *part, *obs, *inst_e;
  int* acc;
  long long* ip64;
  ChainDev* st;
};

struct BSeg {
  int chain, lo, n, start;  // start: first global work index of this chain's slice
};

struct BatchDev {
  unsigned long long exch_seed;
  long long exch_count;
  int parity, pad;
};

constexpr uint32_t SLOT_EXCH = 4u;

__device__ __forceinline__ int find_seg(const BSeg* __restrict__ seg, int nseg, int idx) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg[mid].start <= idx) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int LMAX>
__global__ __launch_bounds__(256) void k_bpropose(const BChain* __restrict__ ch,
                                                  const BSeg* __restrict__ seg, int nseg,
                                                  int total) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const BSeg sg = seg[find_seg(seg, nseg, idx)];
  const int i = idx - sg.start;
  if (i >= sg.n) return;
  const BChain* c = &ch[sg.chain];
  const ChainDev* st = c->st;
  propose_one<LMAX>(c->T, c->config, c->prop, sg.lo + i, st->seed,
                    (int32_t)(st->sweep_index + 1), st->step);
}

template <int L, int LMAX>
__global__ __launch_bounds__(TPB) void k_bmetro(const BChain* __restrict__ ch,
                                                const BSeg* __restrict__ seg, int nseg,
                                                int total_groups, int stride) {
  constexpr int G = TPB / L;
  extern __shared__ double lds[];
  // the slices start at multiples of the wavefront's sites (d_seg_wcol), so
  // the wavefront's first site names the chain of all of them: its tables
  // and buffers are wave-uniform (scalar registers)
  const int gw0 = __builtin_amdgcn_readfirstlane(xcd_swizzle(blockIdx.x, gridDim.x) * G +
                                                 (int)(threadIdx.x / 64) * (64 / L));
  const int gidx = gw0 + (int)(threadIdx.x % 64) / L;
  const int k = __builtin_amdgcn_readfirstlane(find_seg(seg, nseg, gw0 < total_groups ? gw0 : total_groups - 1));
  const BSeg sg = seg[k];
  const int i = gidx - sg.start;
  const bool valid = gidx < total_groups && i < sg.n;
  const BChain* c = &ch[sg.chain];
  metro_group<L, LMAX, false>(c->T, c->config, c->zrows, c->prop, c->dE, c->acc,
                              valid ? c->st->beta : 0.0, sg.lo + i, valid, nullptr, lds, stride);
}

// The acceptance and write-back of one site of chain c at position q (site s)
// by a whole wavefront, as metro_group<64> makes them: accept on
// a <= 0 or u_acc < exp(-beta a), then the proposal's row and spin and the
// position's dE and accept flag.
template <int NLM>
__device__ __forceinline__ void bsite_accept(const BChain* c, int q, int s, double a, double u_acc, double beta,
                                             int lane) {
  constexpr int PS = NLM + 4;
  const double* pq = c->prop + (size_t)q * PS;
  int accept = 0;
  if (lane == 0) accept = (a <= 0.0 || u_acc < sce::det_exp(-beta * a)) ? 1 : 0;
  accept = __shfl(accept, 0, 64);
  if (accept) {
    for (int m = lane; m < NLM; m += 64) c->zrows[(size_t)s * NLM + m] = pq[4 + m];
    if (lane == 0) {
      c->config[3 * (size_t)s] = pq[0];
      c->config[3 * (size_t)s + 1] = pq[1];
      c->config[3 * (size_t)s + 2] = pq[2];
    }
  }
  if (lane == 0) {
    c->dE[q] = accept ? a : 0.0;
    c->acc[q] = accept;
  }
}

// The walk of walk_delta<64> for K sites that share one sublattice walk, each
// with its own staged block (LDS byte bases zb[k]): every entry is loaded once
// and used K times, and site k's sum is the one walk_delta<64> forms for it
// (the same operations in the same order on the same values).
template <int K, int NLM>
__device__ __forceinline__ void walk_delta_shared(const WalkEntry* __restrict__ wl, int Wp, int2 pre, int lane,
                                                  const unsigned (&zb)[K], double (&a)[K]) {
  constexpr int L = 64, B = WALK_UNROLL * L;
  const WalkWords* __restrict__ ww = reinterpret_cast<const WalkWords*>(wl) + lane;
  const int n1 = pre.x - pre.x % B, n2 = pre.y - pre.y % B;
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = 0.0;
  int base = 0;
  for (; base < n1; base += B) {
    WalkWords en[WALK_UNROLL];
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u) en[u] = ld_walk(ww + base + u * L);
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double p = lds_at(lds_add_lo(zb[k], en[u].f01));
        a[k] = __builtin_fma(en[u].w * p, lds_at(lds_add_hi(zb[k], en[u].f2t)), a[k]);
      }
  }
  for (; base < n2; base += B) {
    WalkWords en[WALK_UNROLL];
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u) en[u] = ld_walk(ww + base + u * L);
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double p = lds_at(lds_add_lo(zb[k], en[u].f01)) * lds_at(lds_add_hi(zb[k], en[u].f01));
        a[k] = __builtin_fma(en[u].w * p, lds_at(lds_add_hi(zb[k], en[u].f2t)), a[k]);
      }
  }
  for (; base < Wp; base += B) {
    WalkWords en[WALK_UNROLL];
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u) en[u] = ld_walk(ww + base + u * L);
#pragma unroll
    for (int u = 0; u < WALK_UNROLL; ++u)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double p = (lds_at(lds_add_lo(zb[k], en[u].f01)) * lds_at(lds_add_hi(zb[k], en[u].f01))) *
                         lds_at(lds_add_lo(zb[k], en[u].f2t));
        a[k] = __builtin_fma(en[u].w * p, lds_at(lds_add_hi(zb[k], en[u].f2t)), a[k]);
      }
  }
}

// One colour step of a ladder whose chains share one model (a homogeneous
// batch): wavefront w takes position q = lo + w % n of the colour class for
// the K chains K (w / n) ... K (w / n) + K - 1.  Their sites are the same site
// of the same layout, so they share its sublattice walk: one wavefront walks
// it once for all K, 64 lanes per site.  Each chain's update is the one
// k_bmetro<64> makes (same staging, walk sums, fold and acceptance), so the
// results are those of 64 lanes per site.
template <int K, int LMAX>
__global__ __launch_bounds__(TPB) void k_bmetro_shared(const BChain* __restrict__ ch, int R, int lo, int n,
                                                       int total_waves, int stride) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1), PS = NLM + 4;
  extern __shared__ double lds[];
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x / 64);
  const int gw = __builtin_amdgcn_readfirstlane(xcd_swizzle(blockIdx.x, gridDim.x) * (TPB / 64) + wv);
  const int lane = (int)threadIdx.x % 64;
  const bool wvalid = gw < total_waves;
  const int q = lo + (wvalid ? gw % n : 0);
  const int r0 = (wvalid ? gw / n : 0) * K;
  const DevTables& T = ch[0].T;
  double* Zw = lds + (size_t)wv * K * stride;
  const int dzo = T.nb_max * NLM + 1;
  unsigned zb[K];
#pragma unroll
  for (int k = 0; k < K; ++k) zb[k] = (unsigned)(size_t)(lds_cdouble*)(Zw + (size_t)k * stride);
  int s = 0;
  const WalkEntry* wl = nullptr;
  int W = 0;
  int2 pre = make_int2(0, 0);
  double u_acc[K];
  if (wvalid) {
    s = __builtin_amdgcn_readfirstlane(ldg(T.color_sites + q));
    const int a0 = s % T.na;
    const int w0 = ldg(T.walkp_ptr + a0);
    W = ldg(T.walkp_ptr + a0 + 1) - w0;
    wl = T.walkp + w0;
    pre = ldg_int2(T.walkp_pre + a0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      u_acc[k] = 0.0;
      if (r0 + k >= R) continue;
      const BChain* c = &ch[r0 + k];
      const double* prop = c->prop;
      const double* zr = c->zrows;
      double* Z = Zw + (size_t)k * stride;
      RowDelta<64, NLM> rd;
      u_acc[k] = ldg(prop + (size_t)q * PS + 3);
      stage_neighbours_q<64, NLM>(T, zr, q, s, lane, Z, [&] {
        rd.load(prop + (size_t)q * PS + 4, zr + (size_t)s * NLM, lane);
      });
      rd.store(Z + dzo, lane);
    }
  }
  // the K staged blocks are this wavefront's own
  wave_lds_sync();
  double a[K];
  if (wvalid) walk_delta_shared<K, NLM>(wl, W, pre, lane, zb, a);
  else
#pragma unroll
    for (int k = 0; k < K; ++k) a[k] = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = fold_lanes<64>(a[k]);
  if (!wvalid) return;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (r0 + k >= R) break;
    const BChain* c = &ch[r0 + k];
    bsite_accept<NLM>(c, q, s, a[k], u_acc[k], c->st->beta, lane);
  }
}

// One colour step of any batch with two sites per wavefront: pair w of the
// colour's table is {chain, q1, q2} with q1 and q2 positions of that chain's
// class on one sublattice (q2 = -1: q1 alone), so both sites walk the same
// sublattice walk, once (walk_delta_shared), 64 lanes per site each.  Each
// site's update is the one k_bmetro<64> makes, so the results are those of
// 64 lanes per site.
template <int LMAX>
__global__ __launch_bounds__(TPB) void k_bmetro_pair(const BChain* __restrict__ ch, const int4* __restrict__ pairs,
                                                     int npairs, int stride) {
  constexpr int K = 2, NLM = (LMAX + 1) * (LMAX + 1), PS = NLM + 4;
  extern __shared__ double lds[];
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x / 64);
  const int gw = __builtin_amdgcn_readfirstlane(xcd_swizzle(blockIdx.x, gridDim.x) * (TPB / 64) + wv);
  const int lane = (int)threadIdx.x % 64;
  const bool wvalid = gw < npairs;
  const int4 pr = pairs[wvalid ? gw : 0];
  const int chain = __builtin_amdgcn_readfirstlane(pr.x);
  int q[K] = {__builtin_amdgcn_readfirstlane(pr.y), __builtin_amdgcn_readfirstlane(pr.z)};
  const bool valid[K] = {wvalid, wvalid && q[1] >= 0};
  if (!valid[1]) q[1] = q[0];
  const BChain* c = &ch[chain];
  const DevTables& T = c->T;
  const double* prop = c->prop;
  const double* zr = c->zrows;
  double* Zw = lds + (size_t)wv * K * stride;
  const int dzo = T.nb_max * NLM + 1;
  unsigned zb[K];
#pragma unroll
  for (int k = 0; k < K; ++k) zb[k] = (unsigned)(size_t)(lds_cdouble*)(Zw + (size_t)k * stride);
  int s[K] = {0, 0};
  const WalkEntry* wl = nullptr;
  int W = 0;
  int2 pre = make_int2(0, 0);
  double u_acc[K] = {0.0, 0.0};
  if (wvalid) {
#pragma unroll
    for (int k = 0; k < K; ++k) s[k] = __builtin_amdgcn_readfirstlane(ldg(T.color_sites + q[k]));
    const int a0 = s[0] % T.na;
    const int w0 = ldg(T.walkp_ptr + a0);
    W = ldg(T.walkp_ptr + a0 + 1) - w0;
    wl = T.walkp + w0;
    pre = ldg_int2(T.walkp_pre + a0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!valid[k]) continue;
      double* Z = Zw + (size_t)k * stride;
      RowDelta<64, NLM> rd;
      u_acc[k] = ldg(prop + (size_t)q[k] * PS + 3);
      stage_neighbours_q<64, NLM>(T, zr, q[k], s[k], lane, Z, [&] {
        rd.load(prop + (size_t)q[k] * PS + 4, zr + (size_t)s[k] * NLM, lane);
      });
      rd.store(Z + dzo, lane);
    }
  }
  // the two staged blocks are this wavefront's own
  wave_lds_sync();
  double a[K];
  if (wvalid) walk_delta_shared<K, NLM>(wl, W, pre, lane, zb, a);
  else
#pragma unroll
    for (int k = 0; k < K; ++k) a[k] = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = fold_lanes<64>(a[k]);
  if (!wvalid) return;
  const double beta = c->st->beta;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!valid[k]) break;
    bsite_accept<NLM>(c, q[k], s[k], a[k], u_acc[k], beta, lane);
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_bchunks(const BChain* __restrict__ ch,
                                                 const BSeg* __restrict__ seg, int nseg,
                                                 int meas) {
  const BSeg sg = seg[find_seg(seg, nseg, (int)blockIdx.x)];
  const long long local = (long long)blockIdx.x - sg.start;
  const BChain* c = &ch[sg.chain];
  const long long n = MODE == MODE_MEASURE ? 0 : c->T.n_active;
  chunk_body<MODE>(local, c->dE, n, c->acc, nullptr, c->part, c->ip64, c->config, c->T.active,
                   c->T.n_sites, meas);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_bfinal(const BChain* __restrict__ ch, int meas) {
  const BChain* c = &ch[blockIdx.x];
  const long long n = MODE == MODE_MEASURE ? 0 : c->T.n_active;
  const long long nbx = std::max<long long>((n + TREE_B - 1) / TREE_B, 1);
  const long long nchm = meas ? ((long long)c->T.n_sites + TREE_B - 1) / TREE_B : 0;
  double* scratch = c->part + nbx + 3 * std::max<long long>(nchm, 1);
  final_body<MODE>(c->part, n, c->ip64, c->T.n_sites, meas, c->st, c->obs, c->T.n_active,
                   (double)c->T.n_active, scratch);
}

// Adjacent-pair exchange on the device (gpu_pt.jl:115-140 schedule: pairs
// parity, parity+2, ... each with one keyed uniform, then the parity flips).
// The pairs of one parity share no chain, so thread j of the block takes
// pairs parity + 2 j, parity + 2 (j + blockDim), ...; one block.
__global__ void k_exchange(BChain* ch, int R, BatchDev* bd, const double* __restrict__ kts,
                           long long* sw_att, long long* sw_acc) {
  const long long ex = bd->exch_count + 1;
  const int parity = bd->parity;
  for (int i = parity + 2 * (int)threadIdx.x; i < R - 1; i += 2 * (int)blockDim.x) {
    const sce::u32x4 b = sce::philox_block(bd->exch_seed, i + 1, (int32_t)ex, SLOT_EXCH, 0u);
    const double u = sce::uniform01(b.x, b.y);
    ChainDev* sa = ch[i].st;
    ChainDev* sb = ch[i + 1].st;
    const double logw = (1.0 / kts[i] - 1.0 / kts[i + 1]) * (sa->energy - sb->energy);
    const double m = logw < 0.0 ? logw : 0.0;
    sw_att[i] += 1;
    if (u < sce::det_exp(m)) {
      double* t0 = ch[i].config;
      ch[i].config = ch[i + 1].config;
      ch[i + 1].config = t0;
      double* t1 = ch[i].zrows;
      ch[i].zrows = ch[i + 1].zrows;
      ch[i + 1].zrows = t1;
      const double e = sa->energy;
      sa->energy = sb->energy;
      sb->energy = e;
      sw_acc[i] += 1;
    }
  }
  __syncthreads();  // every thread has read the exchange state
  if (threadIdx.x == 0) {
    bd->parity = 1 - parity;
    bd->exch_count = ex;
  }
}

__global__ void k_breset_obs(BChain* ch, int R) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < R) ch[k].st->n_obs = 0;
}

// ---------------------------------------------------------------------------
// Batched dependency-driven sweeps: every chain of a batch (replica ladder or
// uneven chains) in one cooperative launch per segment of sweeps, replica
// exchange included.  Chain k owns waves [wave_lo, wave_lo + wave_n); within
// a chain the schedule is the single-chain one (wave w of W owns positions
// w, w + W, ... and waits on its neighbours' done flags, here counted in
// segment sweeps from 0).  Across chains the only coupling is the exchange:
//  - each wave adds its count of completed positions of (chain, sweep) to
//    cnt; the wave that completes the count reduces the sweep's dE in the
//    chunk and final trees of the per-colour path (so the energy is the same
//    bits) after the chain's previous sweep is reduced (red), and records
//    the energies before and after the sweep;
//  - with an exchange after the sweep, the second reducer of a pair (pcnt)
//    draws the keyed uniform, swaps the payload pointers and energies, and
//    releases both chains (ready); a chain without a partner releases itself;
//  - a wave proposes the next sweep only once its chain is released, from
//    the payload the chain then holds.
// Every wait is on work earlier in (sweep, chain order) and every wave is
// resident, so the launch always drains; waits are bounded anyway (spin
// limit -> *err, every wave leaves).  The sweep's counts, measurements and
// state updates are made after the launch from the segment buffers.
// ---------------------------------------------------------------------------
struct BFlowChain {
  int pos_off, site_off;  // first global colour position and site of the chain
  int wave_lo, wave_n;    // the chain's waves
  int unit_off, n_units;  // the chain's units (runs of <= 64 / L positions of one colour)
  long long part_off, ipart_off;  // per-sweep reduction partials of the chain
};

struct BFlowArgs {
  const BChain* ch;
  const BFlowChain* fc;
  int R, G, NS;   // chains, colour positions and sites of the batch
  int nsweeps;
  long long i0;   // sweeps of the phase before the launch (measurement schedule)
  int measure_interval, defer_last;
  unsigned long long exch_mask[2];  // bit i - 1: an exchange follows segment sweep i
  const BatchDev* bd;
  const double* kts;
  long long *sw_att, *sw_acc;
  double* dE_seg;    // nsweeps x G, colour-position order
  int* acc_seg;      // nsweeps x G
  double* cfg_seg;   // measured sweeps x 3 x NS (spins after the sweep)
  int* done;         // per site: the last segment sweep it completed
  int *cnt, *pcnt;   // R x nsweeps: positions completed; reducers of pair (k, k + 1) done
  double *Ecur, *Estart;  // R: energy now; R x nsweeps: energy before each sweep
  int *red, *ready;  // R: last sweep reduced; last sweep whose exchange is settled
  unsigned long long *pcfg, *pzr;  // R: the payload each chain holds (exchanged in-kernel)
  double* part;      // per (chain, sweep): chunk partials (dE, then the three spin sums)
  long long* ipart;  // per (chain, sweep): accept counts per chunk
  const int* units;  // per chain, n_units + 1 unit starts (colour positions)
  int stride;        // LDS doubles per site group
  int* err;
  unsigned spin_limit;
};

__device__ __forceinline__ unsigned long long ld_ag(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_ag(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool exch_after(const unsigned long long m[2], int i) {
  return i >= 1 && i <= 128 && ((m[(i - 1) >> 6] >> ((i - 1) & 63)) & 1ull);
}
// exchanges of the segment before sweep i
__device__ __forceinline__ int exch_before(const unsigned long long m[2], int i) {
  const int b = i - 1;  // bits 0 .. b-1
  const unsigned long long lo = b >= 64 ? m[0] : (b > 0 ? m[0] & ((1ull << b) - 1ull) : 0ull);
  const unsigned long long hi = b > 64 ? (b >= 128 ? m[1] : m[1] & ((1ull << (b - 64)) - 1ull)) : 0ull;
  return __popcll(lo) + __popcll(hi);
}

// block_tree (256 threads) by one wavefront: lane l holds threads l + 64 j,
// x0[j] and x1[j] their two operands.  The same additions in the same order;
// the result is valid in lane 0.
__device__ __forceinline__ double wave_block_tree(const double x0[4], const double x1[4]) {
  const double v0 = x0[0] + x1[0], v1 = x0[1] + x1[1], v2 = x0[2] + x1[2], v3 = x0[3] + x1[3];
  double v = (v0 + v2) + (v1 + v3);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_down(v, off, 64);
  return v;
}

// The sum of x[0 .. n) by the chunk trees of chunk_body and the final tree of
// final_body (n <= 512 * 512), by one wavefront; valid in every lane.
__device__ double wave_sweep_sum(const double* x, int n) {
  const int lane = threadIdx.x;
  const int nbx = n > 0 ? (n + TREE_B - 1) / TREE_B : 1;
  double p[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) p[j] = 0.0;
  for (int b = 0; b < nbx; ++b) {
    double x0[4], x1[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i0 = b * TREE_B + lane + 64 * j, i1 = i0 + 256;
      x0[j] = i0 < n ? ld_ag(x + i0) : 0.0;
      x1[j] = i1 < n ? ld_ag(x + i1) : 0.0;
    }
    const double v = __shfl(wave_block_tree(x0, x1), 0, 64);
    if (nbx == 1) return v;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if ((b >> 6) == j && lane == (b & 63)) p[j] = v;
  }
  double x0[4], x1[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    x0[j] = p[j];
    x1[j] = p[j + 4];
  }
  return __shfl(wave_block_tree(x0, x1), 0, 64);
}

// bounded wait until *flag >= need; false (and *err set) past the bound
__device__ __forceinline__ bool flow_wait(const int* flag, int need, int* err, unsigned spin_limit) {
  unsigned spins = 0;
  while (ld_ag(flag) < need) {
    if (++spins > spin_limit) {
      if (threadIdx.x == 0) st_ag(err, 1);
      return false;
    }
    if ((spins & 63u) == 0 && ld_ag(err) != 0) return false;
    if (spins > 4) __builtin_amdgcn_s_sleep(1);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return true;
}

// Reduction of sweep i of chain k, and the exchange that follows it (by the
// pair's second reducer).  One wavefront; false when a wait exceeded its bound.
__device__ bool bflow_reduce(const BFlowArgs& a, int k, int i) {
  const int lane = threadIdx.x;
  const BFlowChain f = a.fc[k];
  const int n = a.ch[k].T.n_active;
  if (!flow_wait(a.red + k, i - 1, a.err, a.spin_limit)) return false;
  const double e0 = ld_ag(a.Ecur + k);
  const double e1 = e0 + wave_sweep_sum(a.dE_seg + (size_t)(i - 1) * a.G + f.pos_off, n);
  if (lane == 0) {
    st_ag(a.Estart + (size_t)k * a.nsweeps + (i - 1), e0);
    st_ag(a.Ecur + k, e1);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0);
  if (lane == 0) st_ag(a.red + k, i);
  if (!exch_after(a.exch_mask, i)) return true;
  const int nex = exch_before(a.exch_mask, i);
  const int parity = a.bd->parity ^ (nex & 1);
  const long long ex = a.bd->exch_count + nex + 1;
  int j = -1;
  if (k < a.R - 1 && ((k - parity) & 1) == 0) j = k;
  else if (k >= 1 && ((k - 1 - parity) & 1) == 0) j = k - 1;
  if (j < 0) {
    __builtin_amdgcn_s_waitcnt(0);
    if (lane == 0) st_ag(a.ready + k, i);
    return true;
  }
  int second = 0;
  if (lane == 0)
    second = __hip_atomic_fetch_add(a.pcnt + (size_t)j * a.nsweeps + (i - 1), 1, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) == 1;
  second = __shfl(second, 0, 64);
  if (!second) return true;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane == 0) {
    // the adjacent-pair rule of k_exchange, pair j of exchange ex
    const sce::u32x4 rb = sce::philox_block(a.bd->exch_seed, j + 1, (int32_t)ex, SLOT_EXCH, 0u);
    const double u = sce::uniform01(rb.x, rb.y);
    const double ea = ld_ag(a.Ecur + j), eb = ld_ag(a.Ecur + j + 1);
    const double logw = (1.0 / a.kts[j] - 1.0 / a.kts[j + 1]) * (ea - eb);
    const double m = logw < 0.0 ? logw : 0.0;
    __hip_atomic_fetch_add(a.sw_att + j, 1ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (u < sce::det_exp(m)) {
      const unsigned long long c0 = ld_ag(a.pcfg + j), c1 = ld_ag(a.pcfg + j + 1);
      const unsigned long long z0 = ld_ag(a.pzr + j), z1 = ld_ag(a.pzr + j + 1);
      st_ag(a.pcfg + j, c1);
      st_ag(a.pcfg + j + 1, c0);
      st_ag(a.pzr + j, z1);
      st_ag(a.pzr + j + 1, z0);
      st_ag(a.Ecur + j, eb);
      st_ag(a.Ecur + j + 1, ea);
      __hip_atomic_fetch_add(a.sw_acc + j, 1ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0);
  if (lane == 0) {
    st_ag(a.ready + j, i);
    st_ag(a.ready + j + 1, i);
  }
  return true;
}

// Waves per SIMD of the batched sweep, forced above the natural count (some
// registers spill; batches are throughput-bound).  64 x fcc 12^3 ladder, ms
// per sweep: 64 lanes 1.79 at six vs 2.00 at four (256 uneven chains: 3.24
// vs 3.43); 32 lanes 1.95 at four vs 2.14 at three, 2.04 at five; 16 lanes
// 2.71 at three, 2.74 at four.
template <int L>
struct BFlowOcc {
  static constexpr int lo = L == 64 ? 6 : L == 32 ? 4 : 1, hi = L == 64 ? 6 : L == 32 ? 4 : 8;
};

// Wave w of chain k (group wl of W) owns the chain's units wl, wl + W, ...;
// group g of L lanes takes position g of each.
template <int LMAX, int L>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(BFlowOcc<L>::lo, BFlowOcc<L>::hi)))
void k_bflow(BFlowArgs a) {
  constexpr int NLM = (LMAX + 1) * (LMAX + 1), G = 64 / L, NSL = 64 / L;
  extern __shared__ double lds[];
  const int lane = threadIdx.x, g = lane / L, gl = lane % L;
  // this wave's chain: the last with wave_lo <= blockIdx.x
  int k = 0;
  {
    int lo = 0, hi = a.R - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.fc[mid].wave_lo <= (int)blockIdx.x) lo = mid;
      else hi = mid - 1;
    }
    k = lo;
  }
  const BFlowChain f = a.fc[k];
  const int wl = (int)blockIdx.x - f.wave_lo, W = f.wave_n;
  const BChain* C = a.ch + k;
  const DevTables& T = C->T;
  const int n = T.n_active, nbm = T.nb_max;
  const int* ulo = a.units + f.unit_off;
  const int nu = f.n_units;
  // a chain without active sites has one wave, which only reduces (to 0)
  if (wl >= W || (nu > 0 && wl >= nu)) return;
  double* Z = lds + (size_t)g * a.stride;
  double* dz = Z + nbm * NLM + 1;
  const ChainDev* st = C->st;
  const double beta = st->beta, step = st->step;
  const unsigned long long seed = st->seed;
  const long long base = st->sweep_index;  // sweeps before the launch
  int* done = a.done + f.site_off;
  double* config = reinterpret_cast<double*>(ld_ag(a.pcfg + k));
  double* zrows = reinterpret_cast<double*>(ld_ag(a.pzr + k));
  // the position tables of unit u (group g), prefetched one item ahead
  int q_nx = 0, s_nx = 0, nb_nx[NSL];
  bool v_nx = false;
  unsigned long long low_nx = 0;
  auto tables = [&](int u) {
    const int q0 = ulo[u];
    q_nx = q0 + g;
    v_nx = q_nx < ulo[u + 1];
    if (!v_nx) q_nx = q0;
    s_nx = T.color_sites[q_nx];
    low_nx = T.nbq_low[q_nx];
#pragma unroll
    for (int m = 0; m < NSL; ++m) {
      const int j = gl + m * L;
      nb_nx[m] = j < nbm ? T.nbq[(size_t)q_nx * nbm + j] : -1;
    }
  };
  if (nu > 0) tables(wl);
  for (int i = 1; i <= a.nsweeps; ++i) {
    const int slot = flow_meas_slot(a.i0, i, a.nsweeps, a.measure_interval, a.defer_last);
    int mine = 0;
    for (int u = wl; u < nu; u += W) {
      const int q = q_nx;
      const int s = L == 64 ? __builtin_amdgcn_readfirstlane(s_nx) : s_nx;
      const unsigned long long low = low_nx;
      const bool valid = v_nx;
      int nb[NSL];
#pragma unroll
      for (int m = 0; m < NSL; ++m) nb[m] = nb_nx[m] >= 0 ? nb_nx[m] : s;
      mine += ulo[u + 1] - ulo[u];
      auto next = [&] { tables(u + W < nu ? u + W : wl); };  // wraps to the next sweep's first unit
      const size_t gq = (size_t)(i - 1) * a.G + f.pos_off + q;
      if (!flow_update<LMAX, L>(T, config, zrows, C->prop, done, a.err, a.spin_limit, q, s, low, nb, valid, i, beta,
                                Z, dz, a.dE_seg + gq, a.acc_seg + gq,
                                slot >= 0 ? a.cfg_seg + ((size_t)slot * a.NS + f.site_off + s) * 3 : nullptr,
                                next))
        return;
    }
    // this wave's share of the sweep is published (flow_update waits for its
    // stores before each flag); the wave completing the count reduces
    int last = 0;
    if (lane == 0)
      last = __hip_atomic_fetch_add(a.cnt + (size_t)k * a.nsweeps + (i - 1), mine, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) + mine == n;
    last = __shfl(last, 0, 64);
    if (last) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (!bflow_reduce(a, k, i)) return;
    }
    if (i < a.nsweeps) {
      if (exch_after(a.exch_mask, i)) {
        // the payload this chain holds for the next sweep
        if (!flow_wait(a.ready + k, i, a.err, a.spin_limit)) return;
        config = reinterpret_cast<double*>(ld_ag(a.pcfg + k));
        zrows = reinterpret_cast<double*>(ld_ag(a.pzr + k));
      }
      if (nu > 0) flow_propose_units<LMAX>(T, config, C->prop, ulo, nu, wl, W, G, seed, (int32_t)(base + i + 1), step);
    }
  }
}

// Segment start: done flags (inactive sites never update: always done),
// counters, energies and the payload each chain holds.
__global__ void k_bflow_init(BFlowArgs a) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < a.NS) {
    int lo = 0, hi = a.R - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.fc[mid].site_off <= idx) lo = mid;
      else hi = mid - 1;
    }
    const int s = (int)(idx - a.fc[lo].site_off);
    a.done[idx] = a.ch[lo].T.active[s] ? 0 : 2147483647;
  }
  if (idx < (long long)a.R * a.nsweeps) {
    a.cnt[idx] = 0;
    a.pcnt[idx] = 0;
  }
  if (idx < a.R) {
    const BChain& c = a.ch[idx];
    a.Ecur[idx] = c.st->energy;
    a.red[idx] = 0;
    a.ready[idx] = 0;
    a.pcfg[idx] = reinterpret_cast<unsigned long long>(c.config);
    a.pzr[idx] = reinterpret_cast<unsigned long long>(c.zrows);
  }
}

// first reduction level of every (chain, sweep) of a segment: block (b, i, k)
// is chunk b of sweep i of chain k (chains with fewer chunks return at once)
__global__ __launch_bounds__(256) void k_bflow_chunks(BFlowArgs a) {
  const int b = blockIdx.x, i = blockIdx.y, k = blockIdx.z;
  const BChain& c = a.ch[k];
  const BFlowChain f = a.fc[k];
  const int n = c.T.n_active, N = c.T.n_sites;
  const long long nbx = std::max<long long>((n + TREE_B - 1) / TREE_B, 1), nchm = ((long long)N + TREE_B - 1) / TREE_B;
  const int slot = flow_meas_slot(a.i0, i + 1, a.nsweeps, a.measure_interval, a.defer_last);
  chunk_body<MODE_METRO>(b, a.dE_seg + (size_t)i * a.G + f.pos_off, n, a.acc_seg + (size_t)i * a.G + f.pos_off,
                         nullptr, a.part + f.part_off + (size_t)i * (nbx + 3 * nchm),
                         a.ipart + f.ipart_off + (size_t)i * 2 * nbx,
                         slot >= 0 ? a.cfg_seg + ((size_t)slot * a.NS + f.site_off) * 3 : nullptr, c.T.active, N,
                         slot >= 0 ? 1 : 0);
}

// The segment's final reductions of chain blockIdx.x, in sweep order: each
// starts from the energy the chain held before the sweep (after the previous
// exchange), so energy, counters, sweep index and measurement rows advance
// as in the per-colour path; then the chain takes the energy and payload it
// holds after the segment, and block 0 advances the exchange schedule.
__global__ __launch_bounds__(256) void k_bflow_final(BFlowArgs a, BChain* ch, BatchDev* bd, int nex) {
  const int k = blockIdx.x;
  BChain& c = ch[k];
  const BFlowChain f = a.fc[k];
  const int n = c.T.n_active, N = c.T.n_sites;
  const long long nbx = std::max<long long>((n + TREE_B - 1) / TREE_B, 1), nchm = ((long long)N + TREE_B - 1) / TREE_B;
  double* scratch = c.part + nbx + 3 * std::max<long long>(nchm, 1);
  for (int i = 0; i < a.nsweeps; ++i) {
    const int slot = flow_meas_slot(a.i0, i + 1, a.nsweeps, a.measure_interval, a.defer_last);
    if (threadIdx.x == 0) c.st->energy = a.Estart[(size_t)k * a.nsweeps + i];
    final_body<MODE_METRO>(a.part + f.part_off + (size_t)i * (nbx + 3 * nchm), n,
                           a.ipart + f.ipart_off + (size_t)i * 2 * nbx, N, slot >= 0 ? 1 : 0, c.st, c.obs, n,
                           (double)n, scratch);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    c.st->energy = a.Ecur[k];
    c.config = reinterpret_cast<double*>(a.pcfg[k]);
    c.zrows = reinterpret_cast<double*>(a.pzr[k]);
    if (k == 0) {
      bd->exch_count += nex;
      bd->parity ^= (nex & 1);
    }
  }
}

struct sce_batch {
  int R = 0, lmax = 0, max_colors = 0, stride = 0;
  bool homogeneous = false;
  std::vector<sce_tables*> tabs;
  DevBuf mem;
  BChain* d_ch = nullptr;
  std::vector<BChain> h_ch;
  BSeg* d_seg_prop = nullptr;
  int nseg_prop = 0, total_prop = 0;
  std::vector<BSeg*> d_seg_col;
  std::vector<int> nseg_col, total_col;
  // the same slices laid out for the walk at 16 / 32 / 64 lanes: every slice
  // starts at a multiple of the wavefront's sites (4 / 2 / 1), so that a
  // wavefront's sites belong to one chain
  std::vector<BSeg*> d_seg_wcol[3];
  std::vector<int> nseg_wcol[3], total_wcol[3];
  BSeg* d_seg_chunk[2] = {nullptr, nullptr};  // [no measurement, measurement]
  BSeg* d_seg_meas = nullptr;                 // measurement-only reduction
  int nseg_chunk[2] = {0, 0}, total_chunk[2] = {0, 0}, nseg_meas = 0, total_meas = 0;
  BatchDev* bd = nullptr;
  double* kts = nullptr;
  long long *sw_att = nullptr, *sw_acc = nullptr;
  double* obs = nullptr;
  size_t obs_cap = 0;
  hipStream_t stream = nullptr;
  GraphSlot gexec[4];  // [meas*2 + exch]
  int g_lanes = 0;
  long long sweeps_done = 0;
  bool use_graph = true;
  bool persistent = false;
  // chains per wavefront that share one walk at 64 lanes per site
  // (k_bmetro_shared; homogeneous batches only): 0 = two where the stencil
  // fits (the default), 1 = off, 2 or 4
  int share = 0;
  // site pairs per colour (k_bmetro_pair): two same-sublattice sites of one
  // chain's colour class per wavefront at 64 lanes, {chain, q1, q2 (-1: none)}
  bool pairs = false;
  std::vector<int4*> d_pairs;
  std::vector<int> npairs;
  // buffers of the batched dependency-driven sweep (k_bflow)
  struct BFlow {
    int cap = 0, mcap = 0, grid = 0, G = 0, NS = 0;
    double *dE = nullptr, *cfg = nullptr, *Ecur = nullptr, *Estart = nullptr, *part = nullptr;
    int *acc = nullptr, *done = nullptr, *cnt = nullptr, *pcnt = nullptr, *red = nullptr, *ready = nullptr;
    int* err = nullptr;
    long long* ipart = nullptr;
    unsigned long long *pcfg = nullptr, *pzr = nullptr;
    BFlowChain* fc = nullptr;
    std::vector<BFlowChain> hfc;
    int L = 0;              // lanes per site the unit tables and wave groups are for
    int* units = nullptr;   // per chain, its unit starts (n_units + 1)
    bool dirty = false;     // hfc changed since its upload
  } bf;
  void drop_graphs() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& g : gexec) g.reset();
    g_lanes = 0;
  }
  ~sce_batch() {
    drop_graphs();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

template <class T>
static int32_t upload_segs(DevBuf& mem, const std::vector<T>& v, T** out) {
  return dupload(mem, out, v.data(), v.size());
}

extern "C" int32_t sce_batch_create(int32_t R, sce_tables* const* tables, const double* const* configs,
                                    const double* const* zrows, const double* energies,
                                    const uint64_t* seeds, const double* steps, const double* kts,
                                    uint64_t exch_seed, sce_batch** out) {
  if (!out || !tables || !configs || !zrows || !energies || !seeds || !steps || !kts)
    return sce_fail(SCE_EINVAL, "sce_batch_create: NULL argument");
  *out = nullptr;
  if (R < 1) return sce_fail(SCE_EINVAL, "a batch needs at least one chain");
  auto* b = new sce_batch();
  auto fail = [&](int32_t r) {
    delete b;
    return r;
  };
  b->R = R;
  b->lmax = tables[0]->dev.lmax;
  b->homogeneous = true;
  int32_t r;
  for (int k = 0; k < R; ++k) {
    if (!tables[k] || !configs[k] || !zrows[k]) return fail(sce_fail(SCE_EINVAL, "chain %d: NULL argument", k));
    if (tables[k]->dev.lmax != b->lmax)
      return fail(sce_fail(SCE_EINVAL, "all chains of a batch must share lmax (%d vs %d)", tables[k]->dev.lmax, b->lmax));
    if (!(steps[k] > 0.0)) return fail(sce_fail(SCE_EINVAL, "chain %d: step must be > 0", k));
    if (!(kts[k] > 0.0 && std::isfinite(kts[k]))) return fail(sce_fail(SCE_EINVAL, "chain %d: kT must be finite and > 0", k));
    b->homogeneous = b->homogeneous && tables[k] == tables[0];
    b->tabs.push_back(tables[k]);
  }
  const int nlm = (b->lmax + 1) * (b->lmax + 1);
  b->h_ch.resize(R);
  for (int k = 0; k < R; ++k) {
    const DevTables& D = tables[k]->dev;
    BChain& c = b->h_ch[k];
    c.T = D;
    const size_t N = (size_t)D.n_sites, na = (size_t)std::max(D.n_active, 1);
    if ((r = dupload(b->mem, &c.config, configs[k], 3 * N)) != SCE_OK) return fail(r);
    if ((r = dupload(b->mem, &c.zrows, zrows[k], (size_t)nlm * N)) != SCE_OK) return fail(r);
    if ((r = dalloc(b->mem, &c.prop, na * (nlm + 4))) != SCE_OK) return fail(r);
    if ((r = dalloc(b->mem, &c.dE, na)) != SCE_OK) return fail(r);
    if ((r = dalloc(b->mem, &c.acc, na)) != SCE_OK) return fail(r);
    const long long nmax = std::max<long long>({(long long)D.n_active, (long long)D.n_inst, (long long)D.n_sites});
    if ((r = dalloc(b->mem, &c.part, partial_capacity(nmax) + 4 * ((size_t)nmax / TREE_B + 4))) != SCE_OK) return fail(r);
    if ((r = dalloc(b->mem, &c.ip64, 2 * ((size_t)nmax / TREE_B + 4))) != SCE_OK) return fail(r);
    if ((r = dalloc(b->mem, &c.inst_e, (size_t)std::max(D.n_inst, 1))) != SCE_OK) return fail(r);
    if ((r = dalloc(b->mem, &c.st, 1)) != SCE_OK) return fail(r);
    ChainDev h{};
    h.energy = energies[k];
    h.step = steps[k];
    h.beta = 1.0 / kts[k];
    h.seed = seeds[k];
    if (hipMemcpy(c.st, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) return fail(sce_fail(SCE_EHIP, "state upload failed"));
    c.obs = nullptr;
    b->stride = std::max(b->stride, D.nb_max * nlm + 1 + nlm);
    b->max_colors = std::max(b->max_colors, tables[k]->n_colors);
  }
  // work segments
  std::vector<BSeg> sp;
  int tot = 0;
  for (int k = 0; k < R; ++k) {
    sp.push_back({k, 0, tables[k]->dev.n_active, tot});
    tot += tables[k]->dev.n_active;
  }
  b->nseg_prop = (int)sp.size();
  b->total_prop = tot;
  if ((r = upload_segs(b->mem, sp, &b->d_seg_prop)) != SCE_OK) return fail(r);
  for (int col = 0; col < b->max_colors; ++col) {
    std::vector<BSeg> sc;
    int t = 0;
    for (int k = 0; k < R; ++k) {
      const sce_tables* tb = tables[k];
      if (col >= tb->n_colors) continue;
      const int lo = tb->color_ptr[col], n = tb->color_ptr[col + 1] - lo;
      if (n == 0) continue;
      sc.push_back({k, lo, n, t});
      t += n;
    }
    if (sc.empty()) sc.push_back({0, 0, 0, 0});
    BSeg* d;
    if ((r = upload_segs(b->mem, sc, &d)) != SCE_OK) return fail(r);
    b->d_seg_col.push_back(d);
    b->nseg_col.push_back((int)sc.size());
    b->total_col.push_back(t);
    for (int li = 0; li < 3; ++li) {
      const int gw = 4 >> li;
      std::vector<BSeg> sw;
      int tw = 0;
      for (const BSeg& x : sc) {
        tw = (tw + gw - 1) / gw * gw;
        sw.push_back({x.chain, x.lo, x.n, tw});
        tw += x.n;
      }
      tw = (tw + gw - 1) / gw * gw;
      if ((r = upload_segs(b->mem, sw, &d)) != SCE_OK) return fail(r);
      b->d_seg_wcol[li].push_back(d);
      b->nseg_wcol[li].push_back((int)sw.size());
      b->total_wcol[li].push_back(t == 0 ? 0 : tw);
    }
    // the colour's site pairs: per chain, the class's positions by
    // sublattice, two at a time (a sublattice's odd last one alone)
    std::vector<int4> pr;
    for (int k = 0; k < R; ++k) {
      const sce_tables* tb = tables[k];
      if (col >= tb->n_colors) continue;
      const int lo = tb->color_ptr[col], n = tb->color_ptr[col + 1] - lo, na = tb->dev.na;
      std::vector<std::vector<int>> by((size_t)na);
      for (int q = lo; q < lo + n; ++q) by[(size_t)(tb->color_sites[q] % na)].push_back(q);
      for (const auto& v : by)
        for (size_t i = 0; i < v.size(); i += 2) pr.push_back(make_int4(k, v[i], i + 1 < v.size() ? v[i + 1] : -1, 0));
    }
    b->npairs.push_back((int)pr.size());
    if (pr.empty()) pr.push_back(make_int4(0, 0, -1, 0));
    int4* dp;
    if ((r = dupload(b->mem, &dp, pr.data(), pr.size())) != SCE_OK) return fail(r);
    b->d_pairs.push_back(dp);
  }
  for (int m = 0; m < 3; ++m) {
    std::vector<BSeg> sc;
    int t = 0;
    for (int k = 0; k < R; ++k) {
      const DevTables& D = tables[k]->dev;
      const long long nbx = std::max<long long>((D.n_active + TREE_B - 1) / TREE_B, 1);
      const long long nchm = ((long long)D.n_sites + TREE_B - 1) / TREE_B;
      const long long nb = m == 0 ? nbx : m == 1 ? std::max(nbx, nchm) : nchm;
      sc.push_back({k, 0, (int)nb, t});
      t += (int)nb;
    }
    BSeg* d;
    if ((r = upload_segs(b->mem, sc, &d)) != SCE_OK) return fail(r);
    if (m < 2) {
      b->d_seg_chunk[m] = d;
      b->nseg_chunk[m] = (int)sc.size();
      b->total_chunk[m] = t;
    } else {
      b->d_seg_meas = d;
      b->nseg_meas = (int)sc.size();
      b->total_meas = t;
    }
  }
  BatchDev hb{};
  hb.exch_seed = exch_seed;
  if ((r = dupload(b->mem, &b->bd, &hb, 1)) != SCE_OK) return fail(r);
  if ((r = dupload(b->mem, &b->kts, kts, (size_t)R)) != SCE_OK) return fail(r);
  std::vector<long long> z((size_t)R, 0);
  if ((r = dupload(b->mem, &b->sw_att, z.data(), z.size())) != SCE_OK) return fail(r);
  if ((r = dupload(b->mem, &b->sw_acc, z.data(), z.size())) != SCE_OK) return fail(r);
  if ((r = dupload(b->mem, &b->d_ch, b->h_ch.data(), b->h_ch.size())) != SCE_OK) return fail(r);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(sce_fail(SCE_EHIP, "hipStreamCreate failed"));
  *out = b;
  return SCE_OK;
}

extern "C" int32_t sce_batch_destroy(sce_batch* b) {
  delete b;
  return SCE_OK;
}

template <int LMAX>
static void launch_bpropose(sce_batch* b, hipStream_t s) {
  const int nb = (b->total_prop + 255) / 256;
  if (nb) k_bpropose<LMAX><<<nb, 256, 0, s>>>(b->d_ch, b->d_seg_prop, b->nseg_prop, b->total_prop);
}

template <int L, int LMAX>
static void launch_bmetro_l(sce_batch* b, int col, hipStream_t s) {
  constexpr int G = TPB / L;
  const int li = L == 16 ? 0 : (L == 32 ? 1 : 2);
  const int total = b->total_wcol[li][col];
  const int nb = (total + G - 1) / G;
  const size_t shmem = (size_t)G * (size_t)b->stride * sizeof(double);
  if (nb)
    k_bmetro<L, LMAX><<<nb, TPB, shmem, s>>>(b->d_ch, b->d_seg_wcol[li][col], b->nseg_wcol[li][col], total, b->stride);
}

template <int LMAX>
static void launch_bmetro(int L, sce_batch* b, int col, hipStream_t s) {
  if (L == 16) launch_bmetro_l<16, LMAX>(b, col, s);
  else if (L == 32) launch_bmetro_l<32, LMAX>(b, col, s);
  else launch_bmetro_l<64, LMAX>(b, col, s);
}

// A colour step of a homogeneous batch with K chains per wavefront sharing
// the walk (k_bmetro_shared): position-major over the class, chain groups
// outer.
template <int K, int LMAX>
static void launch_bmetro_shared_k(sce_batch* b, int col, hipStream_t s) {
  const sce_tables* tb = b->tabs[0];
  if (col >= tb->n_colors) return;
  const int lo = tb->color_ptr[col], n = tb->color_ptr[col + 1] - lo;
  if (n == 0) return;
  constexpr int G = TPB / 64;
  const long long waves = (long long)n * ((b->R + K - 1) / K);
  const int nb = (int)((waves + G - 1) / G);
  const size_t shmem = (size_t)G * K * (size_t)b->stride * sizeof(double);
  k_bmetro_shared<K, LMAX><<<nb, TPB, shmem, s>>>(b->d_ch, b->R, lo, n, (int)waves, b->stride);
}

template <int LMAX>
static void launch_bmetro_pair(sce_batch* b, int col, hipStream_t s) {
  constexpr int G = TPB / 64;
  const int n = b->npairs[col];
  if (n == 0) return;
  const size_t shmem = (size_t)G * 2 * (size_t)b->stride * sizeof(double);
  k_bmetro_pair<LMAX><<<(n + G - 1) / G, TPB, shmem, s>>>(b->d_ch, b->d_pairs[col], n, b->stride);
}

template <int LMAX>
static void launch_bmetro_shared(int K, sce_batch* b, int col, hipStream_t s) {
  if (K == 2) launch_bmetro_shared_k<2, LMAX>(b, col, s);
  else launch_bmetro_shared_k<4, LMAX>(b, col, s);
}

static int32_t benqueue_sweep(sce_batch* b, int L, int meas, int exch) {
  if ((size_t)(TPB / L) * (size_t)b->stride * sizeof(double) > 65536)
    return sce_fail(SCE_EINVAL, "neighbour stencil too large for lanes = %d; use more lanes", L);
  const auto shared_lds = [&](int k) { return (size_t)(TPB / 64) * k * (size_t)b->stride * sizeof(double); };
  int share = 1;
  if (L == 64 && b->homogeneous) {
    if (b->share > 1) {
      share = b->share;
      if (shared_lds(share) > 65536)
        return sce_fail(SCE_EINVAL, "neighbour stencil too large for %d chains per shared walk; share fewer", share);
    } else if (b->share == 0 && b->R >= 2 && shared_lds(2) <= 65536) {
      share = 2;  // the default: two chains per walk where the stencil fits
    }
  }
  const bool pairs = L == 64 && share == 1 && b->pairs;
  if (pairs && shared_lds(2) > 65536)
    return sce_fail(SCE_EINVAL, "neighbour stencil too large for site pairs; turn them off");
  hipStream_t s = b->stream;
  SCE_DISPATCH_LMAX(b->lmax, launch_bpropose, b, s);
  for (int col = 0; col < b->max_colors; ++col) {
    if (b->total_col[col] == 0) continue;
    if (share > 1) {
      SCE_DISPATCH_LMAX(b->lmax, launch_bmetro_shared, share, b, col, s);
    } else if (pairs) {
      SCE_DISPATCH_LMAX(b->lmax, launch_bmetro_pair, b, col, s);
    } else {
      SCE_DISPATCH_LMAX(b->lmax, launch_bmetro, L, b, col, s);
    }
  }
  k_bchunks<MODE_METRO><<<b->total_chunk[meas], 256, 0, s>>>(b->d_ch, b->d_seg_chunk[meas], b->nseg_chunk[meas], meas);
  k_bfinal<MODE_METRO><<<b->R, 256, 0, s>>>(b->d_ch, meas);
  if (exch) k_exchange<<<1, 256, 0, s>>>(b->d_ch, b->R, b->bd, b->kts, b->sw_att, b->sw_acc);
  return SCE_OK;
}

static int32_t bensure_graph(sce_batch* b, int L, int meas, int exch) {
  if (b->g_lanes != L) b->drop_graphs();
  GraphSlot& ge = b->gexec[meas * 2 + exch];
  if (ge) return SCE_OK;
  CHK(capture_graph(b->stream, ge, [&] { return benqueue_sweep(b, L, meas, exch); }));
  b->g_lanes = L;
  return SCE_OK;
}

static int32_t bsync_descriptors(sce_batch* b) {
  HIPCHK(hipMemcpyAsync(b->h_ch.data(), b->d_ch, b->h_ch.size() * sizeof(BChain), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return SCE_OK;
}

// Renormalise every chain (per-chain launches; every renorm_interval sweeps).
static int32_t brenorm(sce_batch* b) {
  CHK(bsync_descriptors(b));
  for (int k = 0; k < b->R; ++k) {
    BChain& c = b->h_ch[k];
    const DevTables& D = c.T;
    SCE_DISPATCH_LMAX(D.lmax, launch_renorm, D, c.config, c.zrows, b->stream);
    const int nb = (D.n_inst + 255) / 256;
    if (nb) k_inst_energy<<<nb, 256, 0, b->stream>>>(D, c.zrows, c.inst_e);
    launch_reduce<MODE_RENORM>(c.inst_e, D.n_inst, nullptr, nullptr, c.part, c.ip64, c.config, D.active,
                               D.n_sites, 0, c.st, nullptr, 0, 1.0, b->stream);
  }
  return SCE_OK;
}

// Chains of the batched persistent sweep: position and site offsets, units
// (runs of at most 64 / L consecutive positions of one colour class, one per
// wave and item) and waves per chain.  Every chain gets one wave, then waves
// go one at a time to the chain with the most units per wave, up to what can
// be resident.
template <int LMAX, int L>
static int32_t bflow_resident(sce_batch* b, int* max_waves) {
  const size_t shmem = (size_t)(64 / L) * (size_t)b->stride * sizeof(double);
  const void* f = reinterpret_cast<const void*>(&k_bflow<LMAX, L>);
  int per_cu = 0, cus = 0;
  CHK(resident_per_cu(f, shmem, &per_cu, &cus));
  *max_waves = per_cu * cus;
  return SCE_OK;
}

template <int LMAX>
static int32_t bflow_waves(sce_batch* b, int L, int* max_waves) {
  if (L == 64) return bflow_resident<LMAX, 64>(b, max_waves);
  if constexpr ((LMAX + 1) * (LMAX + 1) <= 32)
    if (L == 32) return bflow_resident<LMAX, 32>(b, max_waves);
  if constexpr ((LMAX + 1) * (LMAX + 1) <= 16)
    if (L == 16) return bflow_resident<LMAX, 16>(b, max_waves);
  return sce_fail(SCE_EINVAL, "persistent batched sweeps at %d lanes per site need (lmax + 1)^2 <= %d; lmax is %d",
                  L, L, LMAX);
}

static int32_t bflow_setup(sce_batch* b, int L) {
  sce_batch::BFlow& F = b->bf;
  if (F.hfc.empty()) {
    for (int k = 0; k < b->R; ++k) {
      const DevTables& D = b->tabs[k]->dev;
      if (!D.nbq_low)
        return sce_fail(SCE_EINVAL, "persistent sweeps need at most 64 neighbour slots; chain %d's stencil has %d", k,
                        D.nb_max);
      if (D.n_active > TREE_B * TREE_B || D.n_sites > TREE_B * TREE_B)
        return sce_fail(SCE_EDIM, "persistent batched sweeps take chains of at most %d sites", TREE_B * TREE_B);
    }
    F.hfc.assign((size_t)b->R, BFlowChain{});
    int pos = 0, site = 0;
    for (int k = 0; k < b->R; ++k) {
      F.hfc[k].pos_off = pos;
      F.hfc[k].site_off = site;
      pos += b->tabs[k]->dev.n_active;
      site += b->tabs[k]->dev.n_sites;
    }
    F.G = pos;
    F.NS = site;
    CHK(dalloc(b->mem, &F.done, (size_t)std::max(site, 1)));
    CHK(dalloc(b->mem, &F.Ecur, (size_t)b->R));
    CHK(dalloc(b->mem, &F.red, (size_t)b->R));
    CHK(dalloc(b->mem, &F.ready, (size_t)b->R));
    CHK(dalloc(b->mem, &F.pcfg, (size_t)b->R));
    CHK(dalloc(b->mem, &F.pzr, (size_t)b->R));
    CHK(dalloc(b->mem, &F.err, 4));
    CHK(dalloc(b->mem, &F.fc, (size_t)b->R));
    F.dirty = true;
  }
  if (F.L == L) return SCE_OK;
  int wmax = 0;
  int32_t r = SCE_OK;
  SCE_DISPATCH_LMAX_R(r, b->lmax, bflow_waves, b, L, &wmax);
  CHK(r);
  if (wmax < b->R)
    return sce_fail(SCE_EINVAL, "persistent batched sweeps need one resident wave per chain: %d chains, %d waves", b->R,
                    wmax);
  const int G = 64 / L;
  std::vector<int> units;
  std::vector<int> nu((size_t)b->R, 0);
  for (int k = 0; k < b->R; ++k) {
    const sce_tables* tb = b->tabs[k];
    F.hfc[k].unit_off = (int)units.size();
    for (int c = 0; c < tb->n_colors; ++c)
      for (int q = tb->color_ptr[c]; q < tb->color_ptr[c + 1]; q += G) units.push_back(q);
    nu[k] = (int)units.size() - F.hfc[k].unit_off;
    units.push_back(tb->dev.n_active);
    F.hfc[k].n_units = nu[k];
  }
  std::vector<int> nw((size_t)b->R, 1);
  auto load = [&](int k) { return ((long long)std::max(nu[k], 1) + nw[k] - 1) / nw[k]; };
  std::priority_queue<std::pair<long long, int>> pq;
  for (int k = 0; k < b->R; ++k) pq.push({load(k), -k});
  int used = b->R;
  while (used < wmax && !pq.empty()) {
    const int k = -pq.top().second;
    pq.pop();
    if (nw[k] >= std::max(nu[k], 1)) break;  // the busiest chain has a wave per unit
    ++nw[k];
    ++used;
    pq.push({load(k), -k});
  }
  int wave = 0;
  for (int k = 0; k < b->R; ++k) {
    F.hfc[k].wave_lo = wave;
    F.hfc[k].wave_n = nw[k];
    wave += nw[k];
  }
  F.grid = wave;
  HIPCHK(hipStreamSynchronize(b->stream));  // earlier launches may still read the tables
  b->mem.release(F.units);
  F.units = nullptr;
  CHK(dupload(b->mem, &F.units, units.data(), units.size()));
  F.L = L;
  F.dirty = true;
  return SCE_OK;
}

// segment buffers for `cap` sweeps of which at most `mcap` are measured
static int32_t bflow_reserve(sce_batch* b, int cap, int mcap) {
  sce_batch::BFlow& F = b->bf;
  if (cap > F.cap) {
    HIPCHK(hipStreamSynchronize(b->stream));
    for (void* p : {(void*)F.dE, (void*)F.acc, (void*)F.cnt, (void*)F.pcnt, (void*)F.Estart, (void*)F.part,
                    (void*)F.ipart})
      b->mem.release(p);
    F.dE = nullptr, F.acc = nullptr, F.cnt = nullptr, F.pcnt = nullptr, F.Estart = nullptr, F.part = nullptr;
    F.ipart = nullptr;
    F.cap = 0;
    long long part = 0, ipart = 0;
    for (int k = 0; k < b->R; ++k) {
      const DevTables& D = b->tabs[k]->dev;
      const long long nbx = std::max<long long>((D.n_active + TREE_B - 1) / TREE_B, 1);
      const long long nchm = ((long long)D.n_sites + TREE_B - 1) / TREE_B;
      F.hfc[k].part_off = part;
      F.hfc[k].ipart_off = ipart;
      part += (long long)cap * (nbx + 3 * nchm);
      ipart += (long long)cap * 2 * nbx;
    }
    CHK(dalloc(b->mem, &F.dE, (size_t)cap * (size_t)std::max(F.G, 1)));
    CHK(dalloc(b->mem, &F.acc, (size_t)cap * (size_t)std::max(F.G, 1)));
    CHK(dalloc(b->mem, &F.cnt, (size_t)cap * (size_t)b->R));
    CHK(dalloc(b->mem, &F.pcnt, (size_t)cap * (size_t)b->R));
    CHK(dalloc(b->mem, &F.Estart, (size_t)cap * (size_t)b->R));
    CHK(dalloc(b->mem, &F.part, (size_t)part));
    CHK(dalloc(b->mem, &F.ipart, (size_t)ipart));
    F.cap = cap;
    F.dirty = true;
  }
  if (mcap > F.mcap) {
    HIPCHK(hipStreamSynchronize(b->stream));
    b->mem.release(F.cfg);
    F.cfg = nullptr;
    F.mcap = 0;
    CHK(dalloc(b->mem, &F.cfg, (size_t)mcap * 3 * (size_t)std::max(F.NS, 1)));
    F.mcap = mcap;
  }
  if (F.dirty) {
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(F.fc, F.hfc.data(), F.hfc.size() * sizeof(BFlowChain), hipMemcpyHostToDevice));
    F.dirty = false;
  }
  return SCE_OK;
}

template <int LMAX, int L>
static int32_t launch_bflow_l(sce_batch* b, BFlowArgs& fa) {
  const void* f = reinterpret_cast<const void*>(&k_bflow<LMAX, L>);
  void* args[] = {&fa};
  HIPCHK(hipLaunchCooperativeKernel(f, dim3((unsigned)b->bf.grid), dim3(64), args,
                                    (unsigned)((size_t)(64 / L) * b->stride * sizeof(double)), b->stream));
  return SCE_OK;
}

template <int LMAX>
static int32_t launch_bflow(sce_batch* b, BFlowArgs& fa) {
  const int L = b->bf.L;
  if (L == 64) return launch_bflow_l<LMAX, 64>(b, fa);
  if constexpr ((LMAX + 1) * (LMAX + 1) <= 32)
    if (L == 32) return launch_bflow_l<LMAX, 32>(b, fa);
  if constexpr ((LMAX + 1) * (LMAX + 1) <= 16)
    if (L == 16) return launch_bflow_l<LMAX, 16>(b, fa);
  return sce_fail(SCE_EINVAL, "persistent batched sweeps: no kernel for %d lanes at lmax %d", L, LMAX);
}

// n sweeps of every chain (sweeps i0 + 1 .. i0 + n of the phase) as one
// batched dependency-driven launch plus the segment's reductions.  mask: the
// segment sweeps followed by an exchange (nex of them).
static int32_t benqueue_flow(sce_batch* b, int n, long long i0, int64_t measure_interval, int defer_last,
                             const unsigned long long mask[2], int nex) {
  if (n <= 0) return SCE_OK;
  const int mi = (int)std::max<int64_t>(measure_interval, 0);
  int nmeas = 0;
  for (int i = 1; i <= n; ++i) nmeas += flow_meas_slot(i0, i, n, mi, defer_last) >= 0 ? 1 : 0;
  CHK(bflow_reserve(b, n, nmeas));
  const sce_batch::BFlow& F = b->bf;
  hipStream_t s = b->stream;
  // first sweep's proposals (the per-colour path's launch)
  SCE_DISPATCH_LMAX(b->lmax, launch_bpropose, b, s);
  BFlowArgs fa{};
  fa.ch = b->d_ch;
  fa.fc = F.fc;
  fa.R = b->R;
  fa.G = F.G;
  fa.NS = F.NS;
  fa.nsweeps = n;
  fa.i0 = i0;
  fa.measure_interval = mi;
  fa.defer_last = defer_last;
  fa.exch_mask[0] = mask[0];
  fa.exch_mask[1] = mask[1];
  fa.bd = b->bd;
  fa.kts = b->kts;
  fa.sw_att = b->sw_att;
  fa.sw_acc = b->sw_acc;
  fa.dE_seg = F.dE;
  fa.acc_seg = F.acc;
  fa.cfg_seg = F.cfg;
  fa.done = F.done;
  fa.cnt = F.cnt;
  fa.pcnt = F.pcnt;
  fa.Ecur = F.Ecur;
  fa.Estart = F.Estart;
  fa.red = F.red;
  fa.ready = F.ready;
  fa.pcfg = F.pcfg;
  fa.pzr = F.pzr;
  fa.part = F.part;
  fa.ipart = F.ipart;
  fa.units = F.units;
  fa.stride = b->stride;
  fa.err = F.err;
  fa.spin_limit = FLOW_SPIN_LIMIT;
  const long long init = std::max<long long>({(long long)F.NS, (long long)b->R * n, 1});
  k_bflow_init<<<(unsigned)((init + 255) / 256), 256, 0, s>>>(fa);
  int32_t r = SCE_OK;
  SCE_DISPATCH_LMAX_R(r, b->lmax, launch_bflow, b, fa);
  CHK(r);
  // the reductions of every (chain, sweep): per-chain chunk counts, padded to
  // the largest chain (blocks past a chain's chunks return at once)
  long long per = 1;
  for (int k = 0; k < b->R; ++k) {
    const DevTables& D = b->tabs[k]->dev;
    const long long nbx = std::max<long long>((D.n_active + TREE_B - 1) / TREE_B, 1);
    const long long nchm = ((long long)D.n_sites + TREE_B - 1) / TREE_B;
    per = std::max(per, nmeas > 0 ? std::max(nbx, nchm) : nbx);
  }
  k_bflow_chunks<<<dim3((unsigned)per, (unsigned)n, (unsigned)b->R), 256, 0, s>>>(fa);
  k_bflow_final<<<b->R, 256, 0, s>>>(fa, b->d_ch, b->bd, nex);
  HIPCHK(hipGetLastError());
  return SCE_OK;
}

extern "C" int32_t sce_batch_set_persistent(sce_batch* b, int32_t on) {
  if (!b) return sce_fail(SCE_EINVAL, "NULL batch");
  b->persistent = on != 0;
  return SCE_OK;
}

extern "C" int32_t sce_batch_set_site_pairs(sce_batch* b, int32_t on) {
  if (!b) return sce_fail(SCE_EINVAL, "NULL batch");
  if ((on != 0) != b->pairs) {
    b->drop_graphs();
    b->pairs = on != 0;
  }
  return SCE_OK;
}

extern "C" int32_t sce_batch_set_shared_walk(sce_batch* b, int32_t chains) {
  if (!b) return sce_fail(SCE_EINVAL, "NULL batch");
  if (chains != 0 && chains != 1 && chains != 2 && chains != 4)
    return sce_fail(SCE_EINVAL, "chains per shared walk must be 0 (default), 1, 2 or 4 (got %d)", chains);
  if (chains > 1 && !b->homogeneous)
    return sce_fail(SCE_EINVAL, "a shared walk needs every chain of the batch on the same tables");
  if (chains != b->share) {
    b->drop_graphs();
    b->share = chains;
  }
  return SCE_OK;
}

// A renormalisation, then the measurement and the exchange of its sweep
// (the reference measures after renormalising, run.jl:191-200).
static int32_t brenorm_then(sce_batch* b, bool meas, bool exch) {
  CHK(brenorm(b));
  if (meas) {
    k_bchunks<MODE_MEASURE><<<b->total_meas, 256, 0, b->stream>>>(b->d_ch, b->d_seg_meas, b->nseg_meas, 1);
    k_bfinal<MODE_MEASURE><<<b->R, 256, 0, b->stream>>>(b->d_ch, 1);
  }
  if (exch) k_exchange<<<1, 256, 0, b->stream>>>(b->d_ch, b->R, b->bd, b->kts, b->sw_att, b->sw_acc);
  return SCE_OK;
}

extern "C" int32_t sce_batch_run(sce_batch* b, int64_t nsweeps, int32_t lanes, int64_t renorm_interval,
                                 int64_t measure_interval, int64_t exchange_interval,
                                 int64_t phase_done, int64_t phase_total, double* obs,
                                 int64_t obs_cap, int64_t* n_obs) {
  if (!b) return sce_fail(SCE_EINVAL, "NULL batch");
  CHK(check_lanes(lanes));
  if (nsweeps < 0) return sce_fail(SCE_EINVAL, "nsweeps must be >= 0");
  if (exchange_interval > 0 && !b->homogeneous)
    return sce_fail(SCE_EINVAL, "replica exchange needs every chain on the same tables");
  if (b->sweeps_done + nsweeps >= 2147483647LL - 1)
    return sce_fail(SCE_ERANGE_COUNTER, "sweep_index would exhaust the 32-bit RNG counter word");
  if (phase_done < 0) return sce_fail(SCE_EINVAL, "phase_done must be >= 0");
  const int64_t want =
      measure_interval > 0 ? (phase_done + nsweeps) / measure_interval - phase_done / measure_interval : 0;
  if (want > 0 && (!obs || obs_cap < want))
    return sce_fail(SCE_EDIM, "obs holds %lld rows per chain; %lld needed", (long long)obs_cap, (long long)want);
  if (want > 0 && (size_t)want > b->obs_cap) {
    const size_t cap = obs_rows_for((size_t)want, b->obs_cap);
    double* p;
    CHK(dalloc(b->mem, &p, 4 * cap * (size_t)b->R));
    b->drop_graphs();
    b->mem.release(b->obs);
    b->obs = p;
    b->obs_cap = cap;
    CHK(bsync_descriptors(b));
    for (int k = 0; k < b->R; ++k) b->h_ch[k].obs = p + 4 * cap * (size_t)k;
    HIPCHK(hipMemcpy(b->d_ch, b->h_ch.data(), b->h_ch.size() * sizeof(BChain), hipMemcpyHostToDevice));
  }
  k_breset_obs<<<(b->R + 255) / 256, 256, 0, b->stream>>>(b->d_ch, b->R);
  const bool flow = b->persistent && nsweeps > 0;
  if (flow) {
    CHK(bflow_setup(b, lanes));
    HIPCHK(hipMemsetAsync(b->bf.err, 0, sizeof(int), b->stream));
    // segments end at renormalisation points and hold at most FLOW_SEG sweeps;
    // a renormalisation's sweep has its measurement and exchange after it
    int64_t i = 0;
    while (i < nsweeps) {
      int64_t seg = std::min<int64_t>(nsweeps - i, FLOW_SEG);
      if (renorm_interval > 0) seg = std::min<int64_t>(seg, renorm_interval - ((phase_done + i) % renorm_interval));
      const int64_t jend = phase_done + i + seg;
      const bool renorm_after = renorm_interval > 0 && jend % renorm_interval == 0;
      unsigned long long mask[2] = {0ull, 0ull};
      int nex = 0;
      for (int64_t u = 1; u <= seg; ++u) {
        const int64_t j = phase_done + i + u;
        const bool ex = exchange_interval > 0 && j % exchange_interval == 0 && j < phase_total;
        if (ex && !(renorm_after && u == seg)) {
          mask[(u - 1) >> 6] |= 1ull << ((u - 1) & 63);
          ++nex;
        }
      }
      CHK(benqueue_flow(b, (int)seg, phase_done + i, measure_interval, renorm_after ? 1 : 0, mask, nex));
      b->sweeps_done += seg;
      i += seg;
      if (renorm_after)
        CHK(brenorm_then(b, measure_interval > 0 && jend % measure_interval == 0,
                         exchange_interval > 0 && jend % exchange_interval == 0 && jend < phase_total));
    }
  } else {
    for (int64_t i = 1; i <= nsweeps; ++i) {
      const int64_t j = phase_done + i;  // sweep number within the phase
      const bool renorm = renorm_interval > 0 && j % renorm_interval == 0;
      const bool meas = measure_interval > 0 && j % measure_interval == 0;
      const bool exch = exchange_interval > 0 && j % exchange_interval == 0 && j < phase_total;
      const bool fused_meas = meas && !renorm, fused_exch = exch && !renorm;
      if (b->use_graph) {
        CHK(bensure_graph(b, lanes, fused_meas, fused_exch));
        HIPCHK(hipGraphLaunch(b->gexec[(fused_meas ? 2 : 0) + (fused_exch ? 1 : 0)].x, b->stream));
      } else {
        CHK(benqueue_sweep(b, lanes, fused_meas, fused_exch));
      }
      b->sweeps_done += 1;
      if (renorm) CHK(brenorm_then(b, meas, exch));
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  if (flow) {
    int e = 0;
    HIPCHK(hipMemcpy(&e, b->bf.err, sizeof e, hipMemcpyDeviceToHost));
    if (e) return sce_fail(SCE_EHIP, "persistent batched sweep: a wait exceeded its bound");
  }
  // device rows: b->obs_cap per chain; the caller's: obs_cap per chain
  if (want > 0)
    HIPCHK(hipMemcpy2D(obs, 4 * (size_t)obs_cap * 8, b->obs, 4 * b->obs_cap * 8, 4 * (size_t)want * 8,
                       (size_t)b->R, hipMemcpyDeviceToHost));
  if (n_obs) *n_obs = want;
  return SCE_OK;
}

extern "C" int32_t sce_batch_download(sce_batch* b, int32_t k, double* config, double* zrows, double* energy) {
  if (!b || k < 0 || k >= b->R) return sce_fail(SCE_EINVAL, "bad batch or chain index");
  CHK(bsync_descriptors(b));
  const BChain& c = b->h_ch[k];
  const int nlm = (b->lmax + 1) * (b->lmax + 1);
  if (config) HIPCHK(hipMemcpy(config, c.config, 3 * (size_t)c.T.n_sites * 8, hipMemcpyDeviceToHost));
  if (zrows) HIPCHK(hipMemcpy(zrows, c.zrows, (size_t)nlm * c.T.n_sites * 8, hipMemcpyDeviceToHost));
  ChainDev h;
  HIPCHK(hipMemcpy(&h, c.st, sizeof h, hipMemcpyDeviceToHost));
  if (energy) *energy = h.energy;
  return SCE_OK;
}

extern "C" int32_t sce_batch_info(sce_batch* b, int32_t k, sce_chain_info* o) {
  if (!b || !o || k < 0 || k >= b->R) return sce_fail(SCE_EINVAL, "bad batch or chain index");
  HIPCHK(hipStreamSynchronize(b->stream));
  ChainDev h;
  HIPCHK(hipMemcpy(&h, b->h_ch[k].st, sizeof h, hipMemcpyDeviceToHost));
  o->energy = h.energy;
  o->step = h.step;
  o->max_drift = h.max_drift;
  o->last_drift = h.last_drift;
  o->sweep_index = h.sweep_index;
  o->acc_metro = h.acc_metro;
  o->att_metro = h.att_metro;
  o->acc_or = h.acc_or;
  o->att_or = h.att_or;
  o->seed = h.seed;
  o->acc_stage1 = h.acc_stage1;
  return SCE_OK;
}

extern "C" int32_t sce_batch_set_step(sce_batch* b, int32_t k, double step) {
  if (!b || k < 0 || k >= b->R) return sce_fail(SCE_EINVAL, "bad batch or chain index");
  if (!(step > 0.0)) return sce_fail(SCE_EINVAL, "step must be > 0");
  ChainDev h;
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(&h, b->h_ch[k].st, sizeof h, hipMemcpyDeviceToHost));
  h.step = step;
  HIPCHK(hipMemcpy(b->h_ch[k].st, &h, sizeof h, hipMemcpyHostToDevice));
  return SCE_OK;
}

extern "C" int32_t sce_batch_reset_counters(sce_batch* b) {
  if (!b) return sce_fail(SCE_EINVAL, "NULL batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  for (int k = 0; k < b->R; ++k) {
    ChainDev h;
    HIPCHK(hipMemcpy(&h, b->h_ch[k].st, sizeof h, hipMemcpyDeviceToHost));
    h.acc_metro = h.att_metro = h.acc_or = h.att_or = 0;
    h.max_drift = 0.0;
    HIPCHK(hipMemcpy(b->h_ch[k].st, &h, sizeof h, hipMemcpyHostToDevice));
  }
  return SCE_OK;
}

extern "C" int32_t sce_batch_renormalize(sce_batch* b) {
  if (!b) return sce_fail(SCE_EINVAL, "NULL batch");
  CHK(brenorm(b));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  return SCE_OK;
}

extern "C" int32_t sce_batch_swap_stats(sce_batch* b, int64_t* att, int64_t* acc) {
  if (!b) return sce_fail(SCE_EINVAL, "NULL batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  if (b->R < 2) return SCE_OK;  // no pairs
  // n_chains - 1 entries: pair (k, k+1) at k
  if (att) HIPCHK(hipMemcpy(att, b->sw_att, (size_t)(b->R - 1) * 8, hipMemcpyDeviceToHost));
  if (acc) HIPCHK(hipMemcpy(acc, b->sw_acc, (size_t)(b->R - 1) * 8, hipMemcpyDeviceToHost));
  return SCE_OK;
}

extern "C" int32_t sce_batch_set_counters(sce_batch* b, int32_t k, int64_t sweep_index, int64_t acc_metro,
                                          int64_t att_metro, int64_t acc_or, int64_t att_or, double max_drift) {
  if (!b || k < 0 || k >= b->R) return sce_fail(SCE_EINVAL, "bad batch or chain index");
  if (sweep_index < 0 || sweep_index >= 2147483647LL - 1)
    return sce_fail(SCE_ERANGE_COUNTER, "sweep_index %lld outside the 32-bit sweep counter", (long long)sweep_index);
  if (acc_metro < 0 || att_metro < acc_metro || acc_or < 0 || att_or < acc_or)
    return sce_fail(SCE_EINVAL, "counters must satisfy 0 <= acc <= att");
  if (!(max_drift >= 0.0)) return sce_fail(SCE_EINVAL, "max_drift must be >= 0");
  ChainDev h;
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(&h, b->h_ch[k].st, sizeof h, hipMemcpyDeviceToHost));
  h.sweep_index = sweep_index;
  h.acc_metro = acc_metro;
  h.att_metro = att_metro;
  h.acc_or = acc_or;
  h.att_or = att_or;
  h.max_drift = max_drift;
  HIPCHK(hipMemcpy(b->h_ch[k].st, &h, sizeof h, hipMemcpyHostToDevice));
  b->sweeps_done = std::max<int64_t>(b->sweeps_done, sweep_index);
  return SCE_OK;
}

extern "C" int32_t sce_batch_get_exchange(sce_batch* b, int64_t* exch_count, int32_t* parity) {
  if (!b || !exch_count || !parity) return sce_fail(SCE_EINVAL, "NULL argument");
  BatchDev h;
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(&h, b->bd, sizeof h, hipMemcpyDeviceToHost));
  *exch_count = h.exch_count;
  *parity = h.parity;
  return SCE_OK;
}

extern "C" int32_t sce_batch_set_exchange(sce_batch* b, int64_t exch_count, int32_t parity,
                                          const int64_t* att, const int64_t* acc) {
  if (!b || !att || !acc) return sce_fail(SCE_EINVAL, "NULL argument");
  if (exch_count < 0 || exch_count >= 2147483647LL - 1)
    return sce_fail(SCE_ERANGE_COUNTER, "exchange count %lld outside the 32-bit counter", (long long)exch_count);
  if (parity != 0 && parity != 1) return sce_fail(SCE_EINVAL, "parity must be 0 or 1");
  for (int i = 0; i + 1 < b->R; ++i)
    if (att[i] < 0 || acc[i] < 0 || acc[i] > att[i])
      return sce_fail(SCE_EINVAL, "exchange statistics must satisfy 0 <= acc <= att");
  BatchDev h;
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(&h, b->bd, sizeof h, hipMemcpyDeviceToHost));
  h.exch_count = exch_count;
  h.parity = parity;
  HIPCHK(hipMemcpy(b->bd, &h, sizeof h, hipMemcpyHostToDevice));
  if (b->R > 1) {
    HIPCHK(hipMemcpy(b->sw_att, att, (size_t)(b->R - 1) * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->sw_acc, acc, (size_t)(b->R - 1) * 8, hipMemcpyHostToDevice));
  }
  return SCE_OK;
}

extern "C" int32_t sce_batch_set_graph(sce_batch* b, int32_t on) {
  if (!b) return sce_fail(SCE_EINVAL, "NULL batch");
  b->use_graph = on != 0;
  return SCE_OK;
}

// ===========================================================================
// Learned delayed acceptance.
//
// Stage 1 (Christen & Fox 2005): a surrogate site energy g(phi(e)) — an MLP
// (F -> 64 -> 64 -> 1, tanh, plus a linear path v.phi) of a rotation-invariant
// descriptor phi of the site's spin against its neighbour spins — gives
// dE_s = g(phi(e')) - g(phi(e)), exactly antisymmetric under e <-> e'.  The
// proposal passes with min(1, exp(-beta dE_s)) (uniform: slot 0 words 3-4).
// Stage 2: the exact FP64 walk dE decides with min(1, exp(-beta (dE - dE_s)))
// (uniform: slot 3 words 1-2).  A stage-1 rejection skips the neighbour-row
// staging and the walk.  Detailed balance holds for any surrogate; the exact
// FP64 dE always makes the final decision.
//
// Inference runs one site per wavefront in FP32 (lane j = hidden unit j) with
// a fixed operation order (restated by the oracle); training uses MFMA FP32
// GEMMs over minibatches of harvested (phi(e), phi(e'), exact dE) tuples.
// ===========================================================================
constexpr int SUR_H = 64;

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Inference reads the hidden layers' weights in quad-packed form:
// WQ[(q * 64 + j) * 4 + r] = W[j][4q + r], so lane j's four weights of input quad
// q are one 16-byte load and a wavefront's load of quad q is 1 KiB contiguous.
// Layer 1 is zero padded to SUR_QB quads (16 inputs) per batch pair; the
// descriptor rows in LDS are zero padded alike.  (Batches of 4 quads, padding
// to 32 inputs, ran 3.5% slower per DA sweep: F = 112 then carried 16 zero
// inputs.)
constexpr int SUR_BQ = 2;              // quads per weight batch (8 floats per lane)
constexpr int SUR_QB = 2 * SUR_BQ;     // quads per double batch
constexpr int SUR_FQ = 4 * SUR_QB;     // inputs per double batch (16)

__host__ __device__ inline int sur_fpad(int F) { return (F + SUR_FQ - 1) / SUR_FQ * SUR_FQ; }

typedef int int32x4 __attribute__((ext_vector_type(4)));

struct SurNet {
  const float *W1, *b1, *W2, *b2, *w3, *v;  // W1: H x F row-major, W2: H x H row-major
  const float *W1Q, *W2Q;                    // inference copies, quad-packed (W1Q: sur_fpad(F) inputs)
  // int8 copies in the operand layout of v_mfma_i32_16x16x64_i8 (B operand):
  // Q1[(ks * 4 + c) * 64 + l] holds V1[16 c + l % 16][64 ks + 16 (l / 16) + 0..15]
  // (input index padded with zeros to sur_fpq(F)), Q2[c * 64 + l] likewise for V2;
  // qr1 / qr2: per-unit dequantisation factors t_j / 127; qin: input scales
  const int32x4 *Q1, *Q2;
  const float *qr1, *qr2, *qin;
  int F;
  double scale;
};

__device__ __forceinline__ double dot3(const double* a, const double* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// The descriptor reads each neighbour's spin back from its tesseral row: with
// Z_{1,-1} = N1 y, Z_{1,0} = N1 z, Z_{1,1} = N1 x (N1 = sqrt(3 / 4pi)), the spin
// is (Z[3], Z[1], Z[2]) * (1 / N1), each a double multiply (tables need l >= 1).
// Rows are what the walk stages anyway, so the delayed-acceptance kernel gathers
// every neighbour once for both stages.  Spin component c comes from row
// sur_spin_row(c).
__host__ __device__ constexpr int sur_spin_row(int c) { return c == 0 ? 3 : c; }

struct SpinsE {  // spins staged as NB x 3
  const double* E;
  __device__ __forceinline__ void get(int k, double* v) const {
    v[0] = E[3 * k];
    v[1] = E[3 * k + 1];
    v[2] = E[3 * k + 2];
  }
};
struct SpinsZ {  // rows staged as NB x nlm
  const double* Z;
  int nlm;
  __device__ __forceinline__ void get(int k, double* v) const {
    const double* z = Z + k * nlm;
    v[0] = z[sur_spin_row(0)] * sce::INV_N1;
    v[1] = z[sur_spin_row(1)] * sce::INV_N1;
    v[2] = z[sur_spin_row(2)] * sce::INV_N1;
  }
};

// Spec of descriptor feature f (meaningful for f >= 2 NB): an in-bounds load
// for every f, so a wavefront's loads issue together.
__device__ __forceinline__ int4 sur_spec(const int4* __restrict__ spec, int nspec, int f, int NB) {
  if (nspec == 0) return int4{0, 0, 0, 0};
  const int k = f - 2 * NB;
  return spec[k < 0 ? 0 : (k < nspec ? k : nspec - 1)];
}

// Descriptor feature f (spec q) of the current spin e and the proposed spin e2
// against the neighbour spins `sp`: *a = phi_f(e), *b = phi_f(e2).
template <class S>
__device__ __forceinline__ void sur_feature_pair(int f, int NB, int4 q, const S& sp, const double* e,
                                                 const double* e2, float* a, float* b) {
  double x[3], y[3], z[3];
  if (f < NB) {
    sp.get(f, x);
    *a = (float)dot3(e, x);
    *b = (float)dot3(e2, x);
    return;
  }
  if (f < 2 * NB) {
    sp.get(f - NB, x);
    const double ta = dot3(e, x), tb = dot3(e2, x);
    *a = (float)(ta * ta);
    *b = (float)(tb * tb);
    return;
  }
  if (q.x == 1) {
    sp.get(q.y, x);
    sp.get(q.z, y);
    const double cr[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
    *a = (float)dot3(e, cr);
    *b = (float)dot3(e2, cr);
    return;
  }
  if (q.x == 2) {
    sp.get(q.y, x);
    sp.get(q.z, y);
    sp.get(q.w, z);
    const double yz = dot3(y, z), xz = dot3(x, z), xy = dot3(x, y);
    *a = (float)((dot3(e, x) * yz + dot3(e, y) * xz) + dot3(e, z) * xy);
    *b = (float)((dot3(e2, x) * yz + dot3(e2, y) * xz) + dot3(e2, z) * xy);
    return;
  }
  *a = (float)e[q.y];
  *b = (float)e2[q.y];
}

// Both descriptor rows of one site, phi(e) -> ra and phi(e2) -> rb, over
// f < width (zero for F <= f), two features per lane per round.
template <int L, class S>
__device__ __forceinline__ void sur_rows(int lane, int F, int width, int NB, const int4* __restrict__ spec,
                                         int nspec, const S& sp, const double* e, const double* e2, float* ra,
                                         float* rb) {
  for (int f0 = 0; f0 < width; f0 += 2 * L) {
    const int f1 = f0 + lane, f2 = f1 + L;
    const int4 q1 = sur_spec(spec, nspec, f1, NB), q2 = sur_spec(spec, nspec, f2, NB);
    float a1 = 0.0f, b1 = 0.0f, a2 = 0.0f, b2 = 0.0f;
    if (f1 < F) sur_feature_pair(f1, NB, q1, sp, e, e2, &a1, &b1);
    if (f2 < F) sur_feature_pair(f2, NB, q2, sp, e, e2, &a2, &b2);
    if (f1 < width) ra[f1] = a1, rb[f1] = b1;
    if (f2 < width) ra[f2] = a2, rb[f2] = b2;
  }
}


// Stages the neighbour spins of site s (read back from their rows) into E (NB x 3).
template <int L>
__device__ __forceinline__ void stage_spins(const DevTables& T, const double* __restrict__ zrows, int s,
                                            int lane, double* E, int* NB_out) {
  const int atom = s % T.na;
  const int cell = s / T.na;
  const int c1 = cell % T.d0;
  const int r = cell / T.d0;
  const int c2 = r % T.d1;
  const int c3 = r / T.d1;
  const int nb0 = T.nb_ptr[atom];
  const int NB = T.nb_ptr[atom + 1] - nb0;
#pragma unroll 2
  for (int idx = lane; idx < 3 * NB; idx += L) {
    const int nbi = idx / 3;
    const int4 o = T.nb[nb0 + nbi];
    int x = c1 + o.y, y = c2 + o.z, z = c3 + o.w;
    if (x >= T.d0) x -= T.d0;
    if (y >= T.d1) y -= T.d1;
    if (z >= T.d2) z -= T.d2;
    const int n = o.x + T.na * (x + T.d0 * (y + T.d1 * z));
    E[idx] = zrows[(size_t)n * T.nlm + sur_spin_row(idx - 3 * nbi)] * sce::INV_N1;
  }
  *NB_out = NB;
}

__device__ __forceinline__ void fma_quad(const floatx4 w, const floatx4 a, const floatx4 b, float& zA, float& zB) {
  zA = __builtin_fmaf(a[0], w[0], zA);
  zB = __builtin_fmaf(b[0], w[0], zB);
  zA = __builtin_fmaf(a[1], w[1], zA);
  zB = __builtin_fmaf(b[1], w[1], zB);
  zA = __builtin_fmaf(a[2], w[2], zA);
  zB = __builtin_fmaf(b[2], w[2], zB);
  zA = __builtin_fmaf(a[3], w[3], zA);
  zB = __builtin_fmaf(b[3], w[3], zB);
}

// Pre-activations of hidden unit j (this lane) for two inputs xA, xB (LDS, nq
// quads, nq a multiple of SUR_QB): the fused chains
// z = fma(x[4nq-1], W[j][4nq-1], ... fma(x[0], W[j][0], z)).  wq = WQ + j in
// quads.  Weights arrive in batches of SUR_BQ quads, one batch in flight while
// the previous one is used.
typedef const __attribute__((address_space(1))) floatx4 gfloatx4;
typedef const __attribute__((address_space(1))) float gfloat;

__device__ __forceinline__ void hidden_pair(const floatx4* wq_generic, int nq, const floatx4* xA,
                                            const floatx4* xB, float& zA, float& zB) {
  // the weights are global memory: global loads count on the in-order vector
  // memory counter only, so each batch is waited for separately
  gfloatx4* __restrict__ wq = (gfloatx4*)wq_generic;
  floatx4 p[SUR_BQ], r[SUR_BQ];
#pragma unroll
  for (int u = 0; u < SUR_BQ; ++u) p[u] = wq[u * SUR_H];
  for (int q0 = 0; q0 < nq; q0 += SUR_QB) {
#pragma unroll
    for (int u = 0; u < SUR_BQ; ++u) r[u] = wq[(q0 + SUR_BQ + u) * SUR_H];
#pragma unroll
    for (int u = 0; u < SUR_BQ; ++u) fma_quad(p[u], xA[q0 + u], xB[q0 + u], zA, zB);
    if (q0 + SUR_QB < nq) {
#pragma unroll
      for (int u = 0; u < SUR_BQ; ++u) p[u] = wq[(q0 + SUR_QB + u) * SUR_H];
    }
#pragma unroll
    for (int u = 0; u < SUR_BQ; ++u) fma_quad(r[u], xA[q0 + SUR_BQ + u], xB[q0 + SUR_BQ + u], zA, zB);
  }
}

// The surrogate difference of one site on one wavefront (lane j = hidden unit
// j).  phiA/phiB: the site's descriptor rows in LDS (16-byte aligned), zero
// padded to sur_fpad(F).  Each pre-activation is the fused multiply-add chain
// z = fma(x[K-1], w[K-1], ... fma(x[0], w[0], b)) in ascending input index — the
// value a matrix-core tile (v_mfma_f32_16x16x4f32: four fused multiply-adds in k
// order per instruction) produces, and the order the oracle restates over the
// input padded to a multiple of 4.  The longer zero padding here appends
// fma(+0, +0, z) steps, which return z except that a z of -0 becomes +0; tanh
// maps both zeros to +0, so the hidden activations are the same bits.  The head
// (w3 . h2 + v . phi) keeps separate multiplies and adds and folds over the 64
// lanes.  h1A/h1B belong to the wavefront: it orders its own LDS only.
__device__ double sur_delta(const SurNet& net, const float* phiA, const float* phiB, float* h1A, float* h1B,
                            bool valid) {
  const int j = threadIdx.x % 64;
  float tA = 0.0f, tB = 0.0f;
  if (valid) {
    float zA = ((gfloat*)net.b1)[j], zB = zA;
    hidden_pair(reinterpret_cast<const floatx4*>(net.W1Q) + j, sur_fpad(net.F) / 4,
                reinterpret_cast<const floatx4*>(phiA), reinterpret_cast<const floatx4*>(phiB), zA, zB);
    __asm__ volatile("" : "+v"(zA), "+v"(zB));  // both chains complete before either tanh
    h1A[j] = sce::det_tanhf(zA);
    h1B[j] = sce::det_tanhf(zB);
  }
  wave_lds_sync();  // per-group LDS: the own wavefront only
  if (valid) {
    float zA = ((gfloat*)net.b2)[j], zB = zA;
    hidden_pair(reinterpret_cast<const floatx4*>(net.W2Q) + j, SUR_H / 4, reinterpret_cast<const floatx4*>(h1A),
                reinterpret_cast<const floatx4*>(h1B), zA, zB);
    __asm__ volatile("" : "+v"(zA), "+v"(zB));
    const float w3 = ((gfloat*)net.w3)[j];
    const float oA = w3 * sce::det_tanhf(zA);
    const float oB = w3 * sce::det_tanhf(zB);
    float sA = 0.0f, sB = 0.0f;
    for (int f = j; f < net.F; f += 64) {
      const float v = ((gfloat*)net.v)[f];
      sA = sA + v * phiA[f];
      sB = sB + v * phiB[f];
    }
    tA = oA + sA;
    tB = oB + sB;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    tA = tA + __shfl_down(tA, off, 64);
    tB = tB + __shfl_down(tB, off, 64);
  }
  return ((double)tB - (double)tA) * net.scale;  // valid in lane 0
}

// The surrogate difference of one site on one wavefront with its sublattice's
// int8 network on the matrix cores.  Rows 0 and 1 of the operand tiles are the
// site's quantised phi(e) and phi(e') (lanes read row (l % 16) & 1, so rows
// 2..15 repeat them); four 16-column tiles cover the 64 hidden units.  A
// layer's pre-activation is fmaf((float)acc, r_j, b_j) of the exact integer
// dot product acc (v_mfma_i32_16x16x64_i8 sums int8 products exactly, |acc| <
// 2^24, so no order of accumulation matters); layer 1's activations enter
// layer 2 as rint(127 tanh(z)).  The integer results pass through LDS so that
// every lane evaluates two activations.  The head is sur_delta's.  All weight
// loads of the network are issued together before the first product.  qg: the
// group's LDS (2 FPQ + 128 bytes of operands, 128 int32 results).  No block
// barrier: the wavefront's own LDS only.
// Lane l's share of a sublattice's int8 network: its operands of both layers
// and the scales, biases and head weight of hidden unit l, every load issued
// before the first use.  (Issued before the descriptor is formed, to overlap
// it, they raised the kernel to 100-114 VGPRs: 0.268 against 0.242 ms per DA
// sweep at bcc 32^3.)
struct QNetLane {
  int32x4 w1[2][4], w2[4];
  float r1, b1, r2, b2, w3;
};

__device__ __forceinline__ QNetLane qnet_load(const SurNet& net) {
  const int l = threadIdx.x % 64;
  const int nks = sur_fpq(net.F) / 64;
  const __attribute__((address_space(1))) int32x4* Q1 = (const __attribute__((address_space(1))) int32x4*)net.Q1;
  const __attribute__((address_space(1))) int32x4* Q2 = (const __attribute__((address_space(1))) int32x4*)net.Q2;
  QNetLane w;
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int c = 0; c < 4; ++c) w.w1[ks][c] = Q1[((ks < nks ? ks : 0) * 4 + c) * 64 + l];
#pragma unroll
  for (int c = 0; c < 4; ++c) w.w2[c] = Q2[c * 64 + l];
  w.r1 = ((gfloat*)net.qr1)[l];
  w.b1 = ((gfloat*)net.b1)[l];
  w.r2 = ((gfloat*)net.qr2)[l];
  w.b2 = ((gfloat*)net.b2)[l];
  w.w3 = ((gfloat*)net.w3)[l];
  return w;
}

__device__ double sur_delta_q(const SurNet& net, const QNetLane& W, const float* phiA, const float* phiB,
                              unsigned char* qg, int FPQ, bool valid) {
  const int l = threadIdx.x % 64;
  int8_t* A1 = reinterpret_cast<int8_t*>(qg);
  int8_t* A2 = A1 + 2 * FPQ;
  int* D = reinterpret_cast<int*>(A2 + 128);
  float tA = 0.0f, tB = 0.0f;
  if (valid) {
    const int nks = sur_fpq(net.F) / 64;
    const auto& w1 = W.w1;
    const auto& w2 = W.w2;
    const float r1 = W.r1, b1 = W.b1, r2 = W.r2, b2 = W.b2, w3 = W.w3;
    gfloat* qin = (gfloat*)net.qin;
    for (int f = l; f < 64 * nks; f += 64) {
      float xa = 0.0f, xb = 0.0f;
      if (f < net.F) {
        const float sc = qin[f];
        xa = fminf(127.0f, fmaxf(-127.0f, __builtin_rintf(phiA[f] * sc)));
        xb = fminf(127.0f, fmaxf(-127.0f, __builtin_rintf(phiB[f] * sc)));
      }
      A1[f] = (int8_t)(int)xa;
      A1[FPQ + f] = (int8_t)(int)xb;
    }
    wave_lds_sync();
    int32x4 acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = int32x4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      if (ks < nks) {
        const int32x4 x = *reinterpret_cast<const int32x4*>(A1 + ((l % 16) & 1) * FPQ + 64 * ks + 16 * (l / 16));
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(x, w1[ks][c], acc[c], 0, 0, 0);
      }
    }
    // rows 0 and 1 sit in lanes 0..15 (elements 0 and 1): through LDS, then
    // lane l evaluates unit l of both rows
    if (l < 16) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        D[16 * c + l] = acc[c][0];
        D[64 + 16 * c + l] = acc[c][1];
      }
    }
    wave_lds_sync();
    A2[l] = (int8_t)(int)__builtin_rintf(127.0f * sce::det_tanhf(__builtin_fmaf((float)D[l], r1, b1)));
    A2[64 + l] = (int8_t)(int)__builtin_rintf(127.0f * sce::det_tanhf(__builtin_fmaf((float)D[64 + l], r1, b1)));
    wave_lds_sync();
    const int32x4 x2 = *reinterpret_cast<const int32x4*>(A2 + ((l % 16) & 1) * 64 + 16 * (l / 16));
    int32x4 acc2[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
      acc2[c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(x2, w2[c], int32x4{0, 0, 0, 0}, 0, 0, 0);
    if (l < 16) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        D[16 * c + l] = acc2[c][0];
        D[64 + 16 * c + l] = acc2[c][1];
      }
    }
    wave_lds_sync();
    const float oA = w3 * sce::det_tanhf(__builtin_fmaf((float)D[l], r2, b2));
    const float oB = w3 * sce::det_tanhf(__builtin_fmaf((float)D[64 + l], r2, b2));
    float sA = 0.0f, sB = 0.0f;
    for (int f = l; f < net.F; f += 64) {
      const float v = ((gfloat*)net.v)[f];
      sA = sA + v * phiA[f];
      sB = sB + v * phiB[f];
    }
    tA = oA + sA;
    tB = oB + sB;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    tA = tA + __shfl_down(tA, off, 64);
    tB = tB + __shfl_down(tB, off, 64);
  }
  return ((double)tB - (double)tA) * net.scale;  // valid in lane 0
}

// Bytes of a group's LDS for sur_delta_q with descriptor rows of FPQ (16-byte
// multiple).
__host__ __device__ inline size_t da_q_group_bytes(int FPQ) { return (size_t)2 * FPQ + 128 + 512; }
__host__ __device__ inline size_t da_q_bytes(int FPQ) { return (size_t)(TPB / 64) * da_q_group_bytes(FPQ); }

// LDS of one DA group (doubles): neighbour rows + dz (walk_stride, rounded up to
// even) | phiA, phiB (FP floats each, FP a multiple of SUR_FQ) | h1A, h1B (64 floats
// each).  Every float block starts 16-byte aligned.
__host__ __device__ inline int da_phi_offset(int walk_stride) { return (walk_stride + 1) & ~1; }
__host__ __device__ inline int da_group_stride(int walk_stride, int FP) {
  return da_phi_offset(walk_stride) + FP + SUR_H;
}
__host__ __device__ inline size_t da_lds_bytes(int walk_stride, int FP) {
  return (size_t)(TPB / 64) * da_group_stride(walk_stride, FP) * 8;
}

// Two-stage delayed-acceptance update: one site per wavefront (64 lanes).  The
// neighbour rows are staged once: the descriptor reads the spins from them and
// the stage-1 survivors walk them.
template <int LMAX, bool Q>
__device__ __forceinline__ void da_body(DevTables T, double* config, double* zrows, const double* prop, double* dE,
                                        int* acc, int* acc1, const ChainDev* st, const SurNet* nets, int q_lo,
                                        int q_n, int walk_stride, int FP, int vb, int nvb) {
  constexpr int L = 64, G = TPB / L, NLM = (LMAX + 1) * (LMAX + 1), PS = NLM + 4;
  extern __shared__ double lds[];
  // one site per wavefront: its index, site and sublattice are wave-uniform
  // (scalar registers, scalar loads)
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / L), lane = threadIdx.x % L;
  const int qi = xcd_swizzle(vb, nvb) * G + g;
  const bool valid = qi < q_n;
  const int q = q_lo + qi;
  double* base = lds + g * da_group_stride(walk_stride, FP);
  double* Z = base;
  double* dz = Z + T.nb_max * NLM + 1;
  float* phiA = reinterpret_cast<float*>(base + da_phi_offset(walk_stride));
  float* phiB = phiA + FP;
  float* h1A = phiB + FP;
  float* h1B = h1A + SUR_H;
  int s = 0, atom = 0, NB = 0;
  const double* pq = prop + (size_t)q * PS;
  double e[3] = {0, 0, 0}, e2[3] = {0, 0, 0};
  SurNet net{};
  if (valid) {
    s = __builtin_amdgcn_readfirstlane(T.color_sites[q]);
    atom = s % T.na;
    net = nets[atom];
    e[0] = config[3 * (size_t)s];
    e[1] = config[3 * (size_t)s + 1];
    e[2] = config[3 * (size_t)s + 2];
    e2[0] = pq[0];
    e2[1] = pq[1];
    e2[2] = pq[2];
    RowDelta<L, NLM> rd;
    stage_neighbours_q<L, NLM>(T, zrows, q, s, lane, Z, [&] { rd.load(pq + 4, zrows + (size_t)s * NLM, lane); });
    rd.store(dz, lane);
    NB = T.nb_ptr[atom + 1] - T.nb_ptr[atom];
  }
  wave_lds_sync();  // per-group LDS: the own wavefront only
  if (valid) {
    const int f0 = T.feat_ptr[atom];
    sur_rows<L>(lane, net.F, sur_fpad(net.F), NB, T.feat + f0, T.feat_ptr[atom + 1] - f0, SpinsZ{Z, NLM}, e, e2,
                phiA, phiB);
  }
  wave_lds_sync();  // per-group LDS: the own wavefront only
  double dEs;
  if constexpr (Q)
    dEs = sur_delta_q(net, qnet_load(net), phiA, phiB,
                      reinterpret_cast<unsigned char*>(lds) + da_lds_bytes(walk_stride, FP) +
                          g * da_q_group_bytes(sur_fpq(T.nfeat_max)),
                      sur_fpq(T.nfeat_max), valid);
  else
    dEs = sur_delta(net, phiA, phiB, h1A, h1B, valid);
  int pass = 0;
  if (valid && lane == 0) pass = (dEs <= 0.0 || pq[3] < sce::det_exp(-st->beta * dEs)) ? 1 : 0;
  pass = __shfl(pass, 0, L);
  const bool walk = valid && pass;
  double a = 0.0;
  if (walk) {
    const int w0 = T.walkp_ptr[atom];
    a = walk_delta<L, NLM>(T.walkp + w0, T.walkp_ptr[atom + 1] - w0, T.walkp_pre[atom], lane, Z);
  }
  a = fold_lanes<L>(a);
  if (!valid) return;
  int accept = 0;
  if (lane == 0 && walk) {
    const double d2 = a - dEs;
    const sce::u32x4 b = sce::philox_block(st->seed, s + 1, (int32_t)(st->sweep_index + 1), sce::SLOT_DA2, 0u);
    accept = (d2 <= 0.0 || sce::uniform01(b.x, b.y) < sce::det_exp(-st->beta * d2)) ? 1 : 0;
  }
  accept = __shfl(accept, 0, L);
  if (accept) {
    for (int k = lane; k < NLM; k += L) zrows[(size_t)s * NLM + k] = pq[4 + k];
    if (lane == 0) {
      config[3 * (size_t)s] = pq[0];
      config[3 * (size_t)s + 1] = pq[1];
      config[3 * (size_t)s + 2] = pq[2];
    }
  }
  if (lane == 0) {
    dE[q] = accept ? a : 0.0;
    acc[q] = accept;
    acc1[q] = pass;
  }
}

template <int LMAX, bool Q>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4))) void k_da(
    DevTables T, double* config, double* zrows, const double* prop, double* dE, int* acc, int* acc1,
    const ChainDev* st, const SurNet* nets, int q_lo, int q_n, int walk_stride, int FP) {
  da_body<LMAX, Q>(T, config, zrows, prop, dE, acc, acc1, st, nets, q_lo, q_n, walk_stride, FP, blockIdx.x,
                   gridDim.x);
}

// A delayed-acceptance colour step of a chained sweep: blocks below nwalk run
// k_da's step, the rest make proposals of other classes (r0, then r1), as
// k_metro_prop's extra blocks do.  (The final reduction of the previous sweep
// runs as its own launch before the sweep's walks: carried in this kernel it
// raised the kernel to 136 VGPRs with spills.)
template <int LMAX, bool Q>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4))) void k_da_chain(
    DevTables T, double* config, double* zrows, double* prop, double* dE, int* acc, int* acc1, const ChainDev* st,
    const SurNet* nets, int q_lo, int q_n, int walk_stride, int FP, int nwalk, PropRange r0, PropRange r1) {
  const int b = (int)blockIdx.x;
  if (b < nwalk) {
    da_body<LMAX, Q>(T, config, zrows, prop, dE, acc, acc1, st, nets, q_lo, q_n, walk_stride, FP, b, nwalk);
    return;
  }
  const int pb = b - nwalk, nb0 = prop_blocks(r0.n);
  if (pb < nb0)
    propose_body<LMAX>(T, config, prop, st, r0.lo, r0.n, pb * TPB + (int)threadIdx.x, r0.off);
  else
    propose_body<LMAX>(T, config, prop, st, r1.lo, r1.n, (pb - nb0) * TPB + (int)threadIdx.x, r1.off);
}

// Descriptor rows of a colour slice (for harvesting and tests): one wavefront
// per site, phi(e) and phi(e') written as FP32 rows of width fmax.
template <bool FROM_PROP>
__global__ __launch_bounds__(TPB) void k_features(DevTables T, const double* __restrict__ config,
                                                  const double* __restrict__ zrows,
                                                  const double* __restrict__ prop, int PS,
                                                  const int* __restrict__ sites, const double* __restrict__ e2s,
                                                  int q_lo, int q_n, float* __restrict__ outA,
                                                  float* __restrict__ outB, int8_t* __restrict__ sub) {
  constexpr int L = 64, G = TPB / L;
  extern __shared__ double lds[];
  const int g = threadIdx.x / L, lane = threadIdx.x % L;
  const int qi = blockIdx.x * G + g;
  const bool valid = qi < q_n;
  double* E = lds + g * 3 * T.nb_max;
  int s = 0, NB = 0;
  if (valid) {
    s = FROM_PROP ? T.color_sites[q_lo + qi] : sites[qi];
    stage_spins<L>(T, zrows, s, lane, E, &NB);
  }
  wave_lds_sync();  // per-group LDS: the own wavefront only
  if (!valid) return;
  const SpinsE sp{E};
  const int atom = s % T.na;
  const double e[3] = {config[3 * (size_t)s], config[3 * (size_t)s + 1], config[3 * (size_t)s + 2]};
  const double* p2 = FROM_PROP ? prop + (size_t)(q_lo + qi) * PS : e2s + 3 * (size_t)qi;
  const double e2[3] = {p2[0], p2[1], p2[2]};
  const int f0 = T.feat_ptr[atom], nspec = T.feat_ptr[atom + 1] - f0;
  sur_rows<L>(lane, 2 * NB + nspec, T.nfeat_max, NB, T.feat + f0, nspec, sp, e, e2,
              outA + (size_t)qi * T.nfeat_max, outB + (size_t)qi * T.nfeat_max);
  if (lane == 0 && sub) sub[qi] = (int8_t)atom;
}

// Surrogate dE of given descriptor rows through the inference path (tests).
template <bool Q>
__global__ __launch_bounds__(TPB) void k_sur_eval(const SurNet* nets, int subl, int n, int fmax, int FP,
                                                  const float* A, const float* B, double* out) {
  extern __shared__ double lds[];
  const int g = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int i = blockIdx.x * (TPB / 64) + g;
  const bool valid = i < n;
  float* phiA = reinterpret_cast<float*>(lds) + (size_t)g * (2 * FP + 2 * SUR_H);
  float* phiB = phiA + FP;
  float* h1A = phiB + FP;
  float* h1B = h1A + SUR_H;
  for (int f = lane; f < FP; f += 64) {
    phiA[f] = valid && f < fmax ? A[(size_t)i * fmax + f] : 0.0f;
    phiB[f] = valid && f < fmax ? B[(size_t)i * fmax + f] : 0.0f;
  }
  __syncthreads();
  const SurNet net = nets[subl];
  double d;
  if constexpr (Q)
    d = sur_delta_q(net, qnet_load(net), phiA, phiB,
                    reinterpret_cast<unsigned char*>(lds) + (size_t)(TPB / 64) * (2 * FP + 2 * SUR_H) * 4 +
                        g * da_q_group_bytes(sur_fpq(fmax)),
                    sur_fpq(fmax), valid);
  else
    d = sur_delta(net, phiA, phiB, h1A, h1B, valid);
  if (valid && lane == 0) out[i] = d;
}

// One v_mfma_f32_16x16x4f32 per test (diagnostic): A[0][k] = a[4t + k],
// B[k][0] = b[4t + k], C[0][0] = c[t], all else 0; writes D[0][0].
__global__ void k_test_mfma(const float* a, const float* b, const float* c, float* out, int n) {
  const int t = blockIdx.x;
  if (t >= n) return;
  const int l = threadIdx.x, i = l % 16, k = l / 16;
  const float av = i == 0 ? a[4 * t + k] : 0.0f;
  const float bv = i == 0 ? b[4 * t + k] : 0.0f;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  if (l == 0) acc[0] = c[t];
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
  if (l == 0) out[t] = acc[0];
}

// ---- training: MFMA FP32 GEMMs ---------------------------------------------
// C (M x N, ldc) = op(A) (M x K) . op(B) (K x N) [+ C when accumulate]; each
// 64-thread wavefront computes one 16 x 16 tile with v_mfma_f32_16x16x4f32.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float* __restrict__ A, int lda,
                                              const float* __restrict__ B, int ldb, float* __restrict__ C,
                                              int ldc, int accumulate) {
  const int lane = threadIdx.x % 64;
  const int wave = blockIdx.x * 4 + threadIdx.x / 64;
  const int tiles_n = (N + 15) / 16;
  const int tm = wave / tiles_n, tn = wave % tiles_n;
  if (tm * 16 >= M) return;
  const int m0 = tm * 16, n0 = tn * 16;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  const int ra = m0 + lane % 16, cb = n0 + lane % 16, kk = lane / 16;
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int k = k0 + kk;
    float a = 0.f, b = 0.f;
    if (ra < M && k < K) a = TA ? A[(size_t)k * lda + ra] : A[(size_t)ra * lda + k];
    if (cb < N && k < K) b = TB ? B[(size_t)cb * ldb + k] : B[(size_t)k * ldb + cb];
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = m0 + 4 * (lane / 16) + i, c = n0 + lane % 16;
    if (r < M && c < N) C[(size_t)r * ldc + c] = accumulate ? C[(size_t)r * ldc + c] + acc[i] : acc[i];
  }
}

// Split-K form for long reductions (the weight gradients, K = 2 x batch rows):
// wave w computes tile w % tiles over the K-chunk w / tiles into its own
// partial P[chunk][M][N]; k_sum_parts adds the chunks in ascending order.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_part(int M, int N, int K, const float* __restrict__ A, int lda,
                                                   const float* __restrict__ B, int ldb, float* __restrict__ P,
                                                   int kchunk, int nchunks) {
  const int lane = threadIdx.x % 64;
  const int wave = blockIdx.x * 4 + threadIdx.x / 64;
  const int tiles_n = (N + 15) / 16, tiles = ((M + 15) / 16) * tiles_n;
  if (wave >= tiles * nchunks) return;
  const int chunk = wave / tiles, t = wave % tiles;
  const int m0 = (t / tiles_n) * 16, n0 = (t % tiles_n) * 16;
  const int kb = chunk * kchunk, ke = min(K, kb + kchunk);
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  const int ra = m0 + lane % 16, cb = n0 + lane % 16, kk = lane / 16;
  for (int k0 = kb; k0 < ke; k0 += 4) {
    const int k = k0 + kk;
    float a = 0.f, b = 0.f;
    if (ra < M && k < ke) a = TA ? A[(size_t)k * lda + ra] : A[(size_t)ra * lda + k];
    if (cb < N && k < ke) b = TB ? B[(size_t)cb * ldb + k] : B[(size_t)k * ldb + cb];
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  }
  float* out = P + (size_t)chunk * M * N;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = m0 + 4 * (lane / 16) + i, c = n0 + lane % 16;
    if (r < M && c < N) out[(size_t)r * N + c] = acc[i];
  }
}

__global__ void k_sum_parts(const float* __restrict__ P, int nchunks, int M, int N, float* __restrict__ C, int ldc,
                            int accumulate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * N) return;
  float a = 0.0f;
  for (int k = 0; k < nchunks; ++k) a += P[(size_t)k * M * N + i];
  const int r = i / N, c = i % N;
  C[(size_t)r * ldc + c] = accumulate ? C[(size_t)r * ldc + c] + a : a;
}

constexpr int GEMM_KCHUNK = 512;

static void gemm(bool ta, bool tb, int M, int N, int K, const float* A, int lda, const float* B, int ldb,
                 float* C, int ldc, int accumulate, hipStream_t s, float* part = nullptr) {
  if (part && K > 2 * GEMM_KCHUNK) {
    const int nchunks = (K + GEMM_KCHUNK - 1) / GEMM_KCHUNK;
    const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
    const int nb = (tiles * nchunks + 3) / 4;
    if (ta && !tb) k_gemm_part<true, false><<<nb, 256, 0, s>>>(M, N, K, A, lda, B, ldb, part, GEMM_KCHUNK, nchunks);
    else if (!ta && !tb) k_gemm_part<false, false><<<nb, 256, 0, s>>>(M, N, K, A, lda, B, ldb, part, GEMM_KCHUNK, nchunks);
    else if (!ta && tb) k_gemm_part<false, true><<<nb, 256, 0, s>>>(M, N, K, A, lda, B, ldb, part, GEMM_KCHUNK, nchunks);
    else k_gemm_part<true, true><<<nb, 256, 0, s>>>(M, N, K, A, lda, B, ldb, part, GEMM_KCHUNK, nchunks);
    k_sum_parts<<<(M * N + 255) / 256, 256, 0, s>>>(part, nchunks, M, N, C, ldc, accumulate);
    return;
  }
  const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
  const int nb = (tiles + 3) / 4;
  if (!ta && !tb) k_gemm<false, false><<<nb, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, accumulate);
  else if (ta && !tb) k_gemm<true, false><<<nb, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, accumulate);
  else if (!ta && tb) k_gemm<false, true><<<nb, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, accumulate);
  else k_gemm<true, true><<<nb, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, accumulate);
}

// Z (rows x H) += bias; H = tanh(Z) (in place)
__global__ void k_bias_tanh(float* Z, const float* b, long long rows, int H) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= rows * H) return;
  Z[i] = tanhf(Z[i] + b[i % H]);
}

// g_r = H2_r . w3 + X_r . v; residual and output-gradient of the loss
// mean((scale (g_B - g_A) - y)^2) / scale^2 over B pairs (rows r and B + r).
__global__ void k_head(const float* H2, const float* w3, const float* X, const float* v, int F, int fmax,
                       int Bn, const double* y, double scale, float* dg, double* loss_terms) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= Bn) return;
  double gA = 0.0, gB = 0.0;
  for (int j = 0; j < SUR_H; ++j) {
    gA += (double)H2[(size_t)r * SUR_H + j] * w3[j];
    gB += (double)H2[(size_t)(Bn + r) * SUR_H + j] * w3[j];
  }
  for (int f = 0; f < F; ++f) {
    gA += (double)X[(size_t)r * fmax + f] * v[f];
    gB += (double)X[(size_t)(Bn + r) * fmax + f] * v[f];
  }
  const double res = (gB - gA) - y[r] / scale;
  loss_terms[r] = res * res;
  const float d = (float)(2.0 * res / Bn);
  dg[r] = -d;
  dg[Bn + r] = d;
}

// dZ2 = (dg w3^T) * (1 - H2^2); also dH2 stored nowhere (fused)
__global__ void k_back2(const float* H2, const float* w3, const float* dg, long long rows, float* dZ2) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= rows * SUR_H) return;
  const float h = H2[i];
  dZ2[i] = dg[i / SUR_H] * w3[i % SUR_H] * (1.0f - h * h);
}

__global__ void k_dtanh(float* dZ, const float* Hh, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float h = Hh[i];
  dZ[i] = dZ[i] * (1.0f - h * h);
}

// out[c] = sum_r M[r, c] (+ optional second matrix weighting) — column sums, one block per column
__global__ void k_colsum(const float* Mx, long long rows, int cols, int ld, float* out) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  double a = 0.0;
  for (long long r = threadIdx.x; r < rows; r += 256) a += Mx[(size_t)r * ld + c];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int off = 128; off >= 1; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = (float)sh[0];
}

// grad = sum_r coef[r] * M[r, c]  (GEMV M^T coef), one block per column
__global__ void k_gemv_t(const float* Mx, const float* coef, long long rows, int cols, int ld, float* out) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  double a = 0.0;
  for (long long r = threadIdx.x; r < rows; r += 256) a += (double)Mx[(size_t)r * ld + c] * coef[r];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int off = 128; off >= 1; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = (float)sh[0];
}

__global__ void k_adam(float* p, const float* g, float* m, float* v2, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  m[i] = b1 * m[i] + (1.f - b1) * gi;
  v2[i] = b2 * v2[i] + (1.f - b2) * gi * gi;
  p[i] -= lr * (m[i] / c1) / (sqrtf(v2[i] / c2) + eps);
}

// Gather the minibatch rows idx[0..B) of the dataset: X = [A rows; B rows], y.
__global__ void k_gather(const float* dA, const float* dB, const double* dy, const long long* idx, int Bn,
                         int fmax, float* X, double* y) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= (long long)Bn * fmax) return;
  const int r = (int)(i / fmax), f = (int)(i % fmax);
  const long long row = idx[r];
  X[(size_t)r * fmax + f] = dA[(size_t)row * fmax + f];
  X[(size_t)(Bn + r) * fmax + f] = dB[(size_t)row * fmax + f];
  if (f == 0) y[r] = dy[row];
}

// Quad-packed inference copies (layout in SurNet): W1Q over FQ = sur_fpad(F)
// inputs (zero for f >= F), then W2Q.
__global__ void k_sur_pack(const float* __restrict__ W1, const float* __restrict__ W2, int F, int FQ,
                           float* __restrict__ W1Q, float* __restrict__ W2Q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = (i / 4) % SUR_H, f = 4 * (i / (4 * SUR_H)) + i % 4;
  if (i < SUR_H * FQ) W1Q[i] = f < F ? W1[(size_t)j * F + f] : 0.0f;
  if (i < SUR_H * SUR_H) W2Q[i] = W2[j * SUR_H + f];
}

// ---- surrogate host side ------------------------------------------------------
struct SurSub {
  int F = 0;
  size_t np = 0;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v2 = nullptr;
  float* wQ = nullptr;  // W1Q (H x sur_fpad(F)) then W2Q (H x H), quad-packed
  int8_t* q = nullptr;  // Q1 (64 sur_fpq(F) bytes) then Q2 (4096 bytes)
  float* qf = nullptr;  // qr1 (H), qr2 (H), qin (sur_fpq(F), zero padded)
  std::vector<float> qhost;  // the quantised network as the oracle reads it (sur_qblob)
  double scale = 1.0;
  long long adam_t = 0;
};

struct sce_surrogate {
  sce_tables* t = nullptr;
  int na = 0, fmax = 0;
  std::vector<SurSub> subs;
  SurNet* d_nets = nullptr;
  bool quant = false;  // stage 1 on the int8 networks (sur_delta_q)
  DevBuf mem;
  // harvested dataset: rows of phi(e), phi(e') (fmax floats each), exact dE, sublattice
  float *dA = nullptr, *dB = nullptr;
  double* dy = nullptr;
  int8_t* dsub = nullptr;
  long long n_rows = 0, cap = 0;
  // training scratch for the current batch size
  int bsz = 0;
  float *X = nullptr, *Z1 = nullptr, *Z2 = nullptr, *dg = nullptr, *dZ2 = nullptr, *dH1 = nullptr;
  float* part = nullptr;  // split-K partial tiles of the weight gradients
  double *yb = nullptr, *loss_terms = nullptr;
  long long* idx = nullptr;
  long long idx_cap = 0;
  DevBuf scratch;
  hipStream_t stream = nullptr;
  ~sce_surrogate() {
    if (stream) (void)hipStreamSynchronize(stream), (void)hipStreamDestroy(stream);
  }
};

static size_t sur_np(int F) { return (size_t)SUR_H * F + SUR_H + (size_t)SUR_H * SUR_H + SUR_H + SUR_H + F; }

static int32_t sur_upload_nets(sce_surrogate* s) {
  std::vector<SurNet> h(s->na);
  for (int a = 0; a < s->na; ++a) {
    SurSub& u = s->subs[a];
    SurNet& n = h[a];
    n.F = u.F;
    n.scale = u.scale;
    if (!u.params) continue;
    n.W1 = u.params;
    n.b1 = n.W1 + (size_t)SUR_H * u.F;
    n.W2 = n.b1 + SUR_H;
    n.b2 = n.W2 + SUR_H * SUR_H;
    n.w3 = n.b2 + SUR_H;
    n.v = n.w3 + SUR_H;
    n.W1Q = u.wQ;
    n.W2Q = u.wQ + (size_t)SUR_H * sur_fpad(u.F);
    n.Q1 = reinterpret_cast<const int32x4*>(u.q);
    n.Q2 = reinterpret_cast<const int32x4*>(u.q + (size_t)SUR_H * sur_fpq(u.F));
    n.qr1 = u.qf;
    n.qr2 = u.qf + SUR_H;
    n.qin = u.qf + 2 * SUR_H;
  }
  HIPCHK(hipMemcpy(s->d_nets, h.data(), h.size() * sizeof(SurNet), hipMemcpyHostToDevice));
  return SCE_OK;
}

// The int8 network of one sublattice from its FP32 parameters (host, FP32
// arithmetic in this order, restated by the tests): for each output unit j of
// a layer with input scales s_k, w'_jk = W[j][k] / s_k, t_j = max_k |w'_jk|,
// V[j][k] = rint(w'_jk * (127 / t_j)) and r_j = t_j / 127 (a unit with t_j = 0
// gets V = 0, r = 0).  Layer 2's inputs are rint(127 h): s_k = 127.
// The blob the oracle reads: V1 (H x F), r1 (H), b1 (H), V2 (H x H), r2 (H),
// b2 (H), w3 (H), v (F), qin (F), as floats.
static void sur_quantise_layer(const float* W, int nout, int nin, const float* sin, std::vector<float>& V,
                               std::vector<float>& r) {
  V.assign((size_t)nout * nin, 0.0f);
  r.assign(nout, 0.0f);
  for (int j = 0; j < nout; ++j) {
    float t = 0.0f;
    for (int k = 0; k < nin; ++k) t = std::max(t, std::fabs(W[(size_t)j * nin + k] / sin[k]));
    if (t == 0.0f) continue;
    const float m = 127.0f / t;
    for (int k = 0; k < nin; ++k) {
      float q = std::rint((W[(size_t)j * nin + k] / sin[k]) * m);
      V[(size_t)j * nin + k] = std::min(127.0f, std::max(-127.0f, q));
    }
    r[j] = t / 127.0f;
  }
}

static size_t sur_qblob_len(int F) { return (size_t)SUR_H * F + 5 * SUR_H + (size_t)SUR_H * SUR_H + 2 * (size_t)F; }

static int32_t sur_quantise(sce_surrogate* s, int a) {
  SurSub& u = s->subs[a];
  const int F = u.F, FPQ = sur_fpq(F), H = SUR_H;
  std::vector<float> p(u.np);
  HIPCHK(hipMemcpy(p.data(), u.params, u.np * sizeof(float), hipMemcpyDeviceToHost));
  const float* W1 = p.data();
  const float* b1 = W1 + (size_t)H * F;
  const float* W2 = b1 + H;
  const float* b2 = W2 + (size_t)H * H;
  const float* w3 = b2 + H;
  const float* v = w3 + H;
  const std::vector<float>& qin = s->t->sub_qin[a];
  std::vector<float> V1, r1, V2, r2, s2(H, 127.0f);
  sur_quantise_layer(W1, H, F, qin.data(), V1, r1);
  sur_quantise_layer(W2, H, H, s2.data(), V2, r2);
  u.qhost.clear();
  u.qhost.insert(u.qhost.end(), V1.begin(), V1.end());
  u.qhost.insert(u.qhost.end(), r1.begin(), r1.end());
  u.qhost.insert(u.qhost.end(), b1, b1 + H);
  u.qhost.insert(u.qhost.end(), V2.begin(), V2.end());
  u.qhost.insert(u.qhost.end(), r2.begin(), r2.end());
  u.qhost.insert(u.qhost.end(), b2, b2 + H);
  u.qhost.insert(u.qhost.end(), w3, w3 + H);
  u.qhost.insert(u.qhost.end(), v, v + F);
  u.qhost.insert(u.qhost.end(), qin.begin(), qin.end());
  // operand layout (SurNet)
  std::vector<int8_t> q((size_t)H * FPQ + (size_t)H * H, 0);
  for (int ks = 0; ks < FPQ / 64; ++ks)
    for (int c = 0; c < 4; ++c)
      for (int l = 0; l < 64; ++l)
        for (int b = 0; b < 16; ++b) {
          const int j = 16 * c + l % 16, k = 64 * ks + 16 * (l / 16) + b;
          q[(((size_t)ks * 4 + c) * 64 + l) * 16 + b] = k < F ? (int8_t)V1[(size_t)j * F + k] : 0;
        }
  int8_t* q2 = q.data() + (size_t)H * FPQ;
  for (int c = 0; c < 4; ++c)
    for (int l = 0; l < 64; ++l)
      for (int b = 0; b < 16; ++b) {
        const int j = 16 * c + l % 16, k = 16 * (l / 16) + b;
        q2[((size_t)c * 64 + l) * 16 + b] = (int8_t)V2[(size_t)j * H + k];
      }
  std::vector<float> qf((size_t)2 * H + FPQ, 0.0f);
  std::copy(r1.begin(), r1.end(), qf.begin());
  std::copy(r2.begin(), r2.end(), qf.begin() + H);
  std::copy(qin.begin(), qin.end(), qf.begin() + 2 * H);
  HIPCHK(hipMemcpy(u.q, q.data(), q.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(u.qf, qf.data(), qf.size() * sizeof(float), hipMemcpyHostToDevice));
  return SCE_OK;
}

// Refreshes the inference copies of sublattice `a` after its parameters changed.
static int32_t sur_refresh(sce_surrogate* s, int a) {
  SurSub& u = s->subs[a];
  if (!u.params) return SCE_OK;
  const int FQ = sur_fpad(u.F);
  const int n = std::max(SUR_H * FQ, SUR_H * SUR_H);
  k_sur_pack<<<(n + 255) / 256, 256, 0, s->stream>>>(u.params, u.params + (size_t)SUR_H * u.F + SUR_H, u.F, FQ,
                                                    u.wQ, u.wQ + (size_t)SUR_H * FQ);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  return sur_quantise(s, a);
}

extern "C" int32_t sce_sur_create(sce_tables* t, uint64_t seed, sce_surrogate** out) {
  if (!t || !out) return sce_fail(SCE_EINVAL, "sce_sur_create: NULL argument");
  *out = nullptr;
  if (t->dev.lmax < 1) return sce_fail(SCE_EINVAL, "sce_sur_create: the descriptor needs lmax >= 1");
  auto* s = new sce_surrogate();
  auto fail = [&](int32_t r) {
    delete s;
    return r;
  };
  s->t = t;
  s->na = t->dev.na;
  s->fmax = t->dev.nfeat_max;
  s->subs.resize(s->na);
  uint64_t x = seed;
  auto unif = [&]() {  // splitmix64 -> (0,1)
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * 0x1p-53;
  };
  int32_t r;
  for (int a = 0; a < s->na; ++a) {
    SurSub& u = s->subs[a];
    u.F = t->sub_F[a];
    if (u.F == 0) continue;
    u.np = sur_np(u.F);
    std::vector<float> p(u.np, 0.0f);
    const double a1 = std::sqrt(6.0 / (u.F + SUR_H)), a2 = std::sqrt(6.0 / (2 * SUR_H));
    size_t o = 0;
    for (size_t i = 0; i < (size_t)SUR_H * u.F; ++i) p[o++] = (float)(a1 * (2 * unif() - 1));
    o += SUR_H;
    for (size_t i = 0; i < (size_t)SUR_H * SUR_H; ++i) p[o++] = (float)(a2 * (2 * unif() - 1));
    o += SUR_H;
    for (int i = 0; i < SUR_H; ++i) p[o++] = (float)(0.1 * (2 * unif() - 1));
    if ((r = dupload(s->mem, &u.params, p.data(), p.size())) != SCE_OK) return fail(r);
    std::vector<float> z(u.np, 0.0f);
    if ((r = dupload(s->mem, &u.grads, z.data(), z.size())) != SCE_OK) return fail(r);
    if ((r = dupload(s->mem, &u.m, z.data(), z.size())) != SCE_OK) return fail(r);
    if ((r = dupload(s->mem, &u.v2, z.data(), z.size())) != SCE_OK) return fail(r);
    if ((r = dalloc(s->mem, &u.wQ, (size_t)SUR_H * sur_fpad(u.F) + SUR_H * SUR_H)) != SCE_OK) return fail(r);
    if ((r = dalloc(s->mem, &u.q, (size_t)SUR_H * sur_fpq(u.F) + SUR_H * SUR_H)) != SCE_OK) return fail(r);
    if ((r = dalloc(s->mem, &u.qf, (size_t)2 * SUR_H + sur_fpq(u.F))) != SCE_OK) return fail(r);
  }
  if ((r = dalloc(s->mem, &s->d_nets, (size_t)s->na)) != SCE_OK) return fail(r);
  if ((r = sur_upload_nets(s)) != SCE_OK) return fail(r);
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(sce_fail(SCE_EHIP, "hipStreamCreate failed"));
  for (int a = 0; a < s->na; ++a)
    if ((r = sur_refresh(s, a)) != SCE_OK) return fail(r);
  *out = s;
  return SCE_OK;
}

extern "C" int32_t sce_sur_destroy(sce_surrogate* s) {
  delete s;
  return SCE_OK;
}

extern "C" int32_t sce_sur_info(const sce_surrogate* s, int32_t sub, int32_t* F, int64_t* n_params,
                                double* scale, int64_t* n_rows) {
  if (!s || sub < 0 || sub >= s->na) return sce_fail(SCE_EINVAL, "bad surrogate or sublattice");
  if (F) *F = s->subs[sub].F;
  if (n_params) *n_params = (int64_t)s->subs[sub].np;
  if (scale) *scale = s->subs[sub].scale;
  if (n_rows) *n_rows = s->n_rows;
  return SCE_OK;
}

extern "C" int32_t sce_sur_get_params(sce_surrogate* s, int32_t sub, float* out) {
  if (!s || !out || sub < 0 || sub >= s->na || !s->subs[sub].params) return sce_fail(SCE_EINVAL, "bad sublattice");
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(out, s->subs[sub].params, s->subs[sub].np * 4, hipMemcpyDeviceToHost));
  return SCE_OK;
}

extern "C" int32_t sce_sur_set_params(sce_surrogate* s, int32_t sub, const float* in, double scale) {
  if (!s || !in || sub < 0 || sub >= s->na || !s->subs[sub].params) return sce_fail(SCE_EINVAL, "bad sublattice");
  if (!(scale > 0.0 && std::isfinite(scale))) return sce_fail(SCE_EINVAL, "scale must be finite and > 0");
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(s->subs[sub].params, in, s->subs[sub].np * 4, hipMemcpyHostToDevice));
  s->subs[sub].scale = scale;
  CHK(sur_refresh(s, sub));
  return sur_upload_nets(s);
}

extern "C" int32_t sce_sur_set_int8(sce_surrogate* s, int32_t on) {
  if (!s) return sce_fail(SCE_EINVAL, "NULL surrogate");
  s->quant = on != 0;
  return SCE_OK;
}

extern "C" int32_t sce_sur_get_qparams(sce_surrogate* s, int32_t sub, float* out) {
  if (!s || !out || sub < 0 || sub >= s->na || !s->subs[sub].params) return sce_fail(SCE_EINVAL, "bad sublattice");
  const SurSub& u = s->subs[sub];
  if (u.qhost.size() != sur_qblob_len(u.F)) return sce_fail(SCE_EINVAL, "no quantised network");
  std::copy(u.qhost.begin(), u.qhost.end(), out);
  return SCE_OK;
}

static int32_t sur_reserve(sce_surrogate* s, long long rows) {
  if (rows <= s->cap) return SCE_OK;
  long long cap = std::max(rows, 2 * s->cap);
  float *A, *B;
  double* y;
  int8_t* sb;
  CHK(dalloc(s->mem, &A, (size_t)cap * s->fmax));
  CHK(dalloc(s->mem, &B, (size_t)cap * s->fmax));
  CHK(dalloc(s->mem, &y, (size_t)cap));
  CHK(dalloc(s->mem, &sb, (size_t)cap));
  if (s->n_rows) {
    HIPCHK(hipMemcpy(A, s->dA, (size_t)s->n_rows * s->fmax * 4, hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(B, s->dB, (size_t)s->n_rows * s->fmax * 4, hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(y, s->dy, (size_t)s->n_rows * 8, hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(sb, s->dsub, (size_t)s->n_rows, hipMemcpyDeviceToDevice));
  }
  s->mem.release(s->dA);
  s->mem.release(s->dB);
  s->mem.release(s->dy);
  s->mem.release(s->dsub);
  s->dA = A;
  s->dB = B;
  s->dy = y;
  s->dsub = sb;
  s->cap = cap;
  return SCE_OK;
}

template <int LMAX>
static void launch_harvest_walk(const DevTables& D, double* config, double* zrows, const double* prop, double* dE,
                                int* acc, const ChainDev* st, int lo, int n, double* raw, hipStream_t s) {
  constexpr int G = TPB / 64;
  const int nb = (n + G - 1) / G;
  if (nb)
    k_metro<64, LMAX, false><<<nb, TPB, walk_lds_bytes(D, 64), s>>>(D, config, zrows, prop, dE, acc, st, lo, n,
                                                                    nullptr, raw);
}

// Plain Metropolis sweeps of `c` that also record, for every attempted site,
// (phi(e), phi(e'), exact dE): the surrogate's training tuples.
extern "C" int32_t sce_sur_harvest(sce_surrogate* s, sce_chain* c, double beta, int64_t nsweeps) {
  if (!s || !c || c->t != s->t) return sce_fail(SCE_EINVAL, "harvest: chain and surrogate on different tables");
  CHK(check_lds(c->t->dev, 64));
  CHK(set_beta(c, beta));
  const DevTables& D = c->t->dev;
  CHK(sur_reserve(s, s->n_rows + nsweeps * (long long)D.n_active));
  HIPCHK(hipStreamSynchronize(s->stream));
  const int PS = D.nlm + 4;
  for (int64_t i = 0; i < nsweeps; ++i) {
    if (c->sweeps_done >= 2147483647LL - 1) return sce_fail(SCE_ERANGE_COUNTER, "sweep counter exhausted");
    const long long row0 = s->n_rows;
    SCE_DISPATCH_LMAX(D.lmax, launch_propose, D, c->config, c->prop, c->st, c->stream);
    for (int col = 0; col < c->t->n_colors; ++col) {
      const int lo = c->t->color_ptr[col], n = c->t->color_ptr[col + 1] - lo;
      if (n == 0) continue;
      const int nb = (n + 3) / 4;
      k_features<true><<<nb, TPB, 4 * 3 * (size_t)D.nb_max * 8, c->stream>>>(
          D, c->config, c->zrows, c->prop, PS, nullptr, nullptr, lo, n, s->dA + (size_t)(row0 + lo) * s->fmax,
          s->dB + (size_t)(row0 + lo) * s->fmax, s->dsub + row0 + lo);
      SCE_DISPATCH_LMAX(D.lmax, launch_harvest_walk, D, c->config, c->zrows, c->prop, c->dE, c->acc, c->st, lo, n,
                        s->dy + row0, c->stream);
    }
    launch_reduce<MODE_METRO>(c->dE, D.n_active, c->acc, nullptr, c->partial, c->ip64, c->config, D.active,
                              D.n_sites, 0, c->st, nullptr, D.n_active, (double)D.n_active, c->stream);
    c->sweeps_done += 1;
    s->n_rows += D.n_active;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return SCE_OK;
}

static int32_t sur_scratch(sce_surrogate* s, int bsz) {
  if (s->bsz == bsz) return SCE_OK;
  s->scratch.~DevBuf();
  new (&s->scratch) DevBuf();
  const size_t rows = 2 * (size_t)bsz;
  CHK(dalloc(s->scratch, &s->X, rows * s->fmax));
  CHK(dalloc(s->scratch, &s->Z1, rows * SUR_H));
  CHK(dalloc(s->scratch, &s->Z2, rows * SUR_H));
  CHK(dalloc(s->scratch, &s->dg, rows));
  CHK(dalloc(s->scratch, &s->dZ2, rows * SUR_H));
  CHK(dalloc(s->scratch, &s->dH1, rows * SUR_H));
  CHK(dalloc(s->scratch, &s->yb, (size_t)bsz));
  const size_t nchunks = (rows + GEMM_KCHUNK - 1) / GEMM_KCHUNK;
  CHK(dalloc(s->scratch, &s->part, nchunks * SUR_H * (size_t)std::max(SUR_H, s->fmax)));
  CHK(dalloc(s->scratch, &s->loss_terms, (size_t)bsz));
  s->bsz = bsz;
  return SCE_OK;
}

// Loss and gradient of one minibatch already gathered into s->X / s->yb.
static void sur_grad_batch(sce_surrogate* s, int sub, int bsz) {
  SurSub& u = s->subs[sub];
  const int F = u.F, fm = s->fmax, rows = 2 * bsz;
  hipStream_t st = s->stream;
  float* W1 = u.params;
  float* b1 = W1 + (size_t)SUR_H * F;
  float* W2 = b1 + SUR_H;
  float* b2 = W2 + SUR_H * SUR_H;
  float* w3 = b2 + SUR_H;
  float* v = w3 + SUR_H;
  float* gW1 = u.grads;
  float* gb1 = gW1 + (size_t)SUR_H * F;
  float* gW2 = gb1 + SUR_H;
  float* gb2 = gW2 + SUR_H * SUR_H;
  float* gw3 = gb2 + SUR_H;
  float* gv = gw3 + SUR_H;
  const long long nh = (long long)rows * SUR_H;
  gemm(false, true, rows, SUR_H, F, s->X, fm, W1, F, s->Z1, SUR_H, 0, st);
  k_bias_tanh<<<(unsigned)((nh + 255) / 256), 256, 0, st>>>(s->Z1, b1, rows, SUR_H);
  gemm(false, true, rows, SUR_H, SUR_H, s->Z1, SUR_H, W2, SUR_H, s->Z2, SUR_H, 0, st);
  k_bias_tanh<<<(unsigned)((nh + 255) / 256), 256, 0, st>>>(s->Z2, b2, rows, SUR_H);
  k_head<<<(bsz + 255) / 256, 256, 0, st>>>(s->Z2, w3, s->X, v, F, fm, bsz, s->yb, u.scale, s->dg, s->loss_terms);
  k_gemv_t<<<SUR_H, 256, 0, st>>>(s->Z2, s->dg, rows, SUR_H, SUR_H, gw3);
  k_gemv_t<<<F, 256, 0, st>>>(s->X, s->dg, rows, F, fm, gv);
  k_back2<<<(unsigned)((nh + 255) / 256), 256, 0, st>>>(s->Z2, w3, s->dg, rows, s->dZ2);
  gemm(true, false, SUR_H, SUR_H, rows, s->dZ2, SUR_H, s->Z1, SUR_H, gW2, SUR_H, 0, st, s->part);
  k_colsum<<<SUR_H, 256, 0, st>>>(s->dZ2, rows, SUR_H, SUR_H, gb2);
  gemm(false, false, rows, SUR_H, SUR_H, s->dZ2, SUR_H, W2, SUR_H, s->dH1, SUR_H, 0, st);
  k_dtanh<<<(unsigned)((nh + 255) / 256), 256, 0, st>>>(s->dH1, s->Z1, nh);
  gemm(true, false, SUR_H, F, rows, s->dH1, SUR_H, s->X, fm, gW1, F, 0, st, s->part);
  k_colsum<<<SUR_H, 256, 0, st>>>(s->dH1, rows, SUR_H, SUR_H, gb1);
}

// Train sublattice `sub`'s network: `epochs` passes of Adam over its harvested
// rows in shuffled minibatches of `bsz` pairs.  The output scale is fixed to the
// rms of the rows' exact dE at the first call.  Writes the mean loss (in units
// of scale^2) of every epoch.
extern "C" int32_t sce_sur_train(sce_surrogate* s, int32_t sub, int32_t epochs, int32_t bsz, double lr,
                                 uint64_t seed, double* loss_out) {
  if (!s || sub < 0 || sub >= s->na || !s->subs[sub].params) return sce_fail(SCE_EINVAL, "bad sublattice");
  if (bsz < 16 || epochs < 0) return sce_fail(SCE_EINVAL, "batch must be >= 16 and epochs >= 0");
  SurSub& u = s->subs[sub];
  HIPCHK(hipStreamSynchronize(s->stream));
  std::vector<int8_t> hs((size_t)s->n_rows);
  std::vector<double> hy((size_t)s->n_rows);
  if (s->n_rows) {
    HIPCHK(hipMemcpy(hs.data(), s->dsub, hs.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hy.data(), s->dy, hy.size() * 8, hipMemcpyDeviceToHost));
  }
  std::vector<long long> rows;
  double ss = 0.0;
  for (long long r = 0; r < s->n_rows; ++r)
    if (hs[r] == sub) {
      rows.push_back(r);
      ss += hy[r] * hy[r];
    }
  if ((long long)rows.size() < bsz) return sce_fail(SCE_EINVAL, "sublattice %d has %zu rows; batch %d", sub, rows.size(), bsz);
  if (u.adam_t == 0) {
    u.scale = std::sqrt(ss / rows.size());
    if (!(u.scale > 0.0)) u.scale = 1.0;
    CHK(sur_upload_nets(s));
  }
  CHK(sur_scratch(s, bsz));
  if ((long long)rows.size() > s->idx_cap) {
    s->mem.release(s->idx);
    CHK(dalloc(s->mem, &s->idx, rows.size()));
    s->idx_cap = (long long)rows.size();
  }
  uint64_t x = seed;
  auto next = [&]() {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  const long long nb = (long long)rows.size() / bsz;
  std::vector<double> lt((size_t)bsz);
  for (int ep = 0; ep < epochs; ++ep) {
    for (size_t i = rows.size() - 1; i > 0; --i) std::swap(rows[i], rows[next() % (i + 1)]);
    HIPCHK(hipMemcpyAsync(s->idx, rows.data(), rows.size() * 8, hipMemcpyHostToDevice, s->stream));
    double lsum = 0.0;
    for (long long bi = 0; bi < nb; ++bi) {
      const long long tot = (long long)bsz * s->fmax;
      k_gather<<<(unsigned)((tot + 255) / 256), 256, 0, s->stream>>>(s->dA, s->dB, s->dy, s->idx + bi * bsz, bsz,
                                                                    s->fmax, s->X, s->yb);
      sur_grad_batch(s, sub, bsz);
      u.adam_t += 1;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)u.adam_t), c2 = 1.f - std::pow(b2, (float)u.adam_t);
      k_adam<<<(unsigned)((u.np + 255) / 256), 256, 0, s->stream>>>(u.params, u.grads, u.m, u.v2, (long long)u.np,
                                                                    (float)lr, b1, b2, 1e-8f, c1, c2);
      if (bi == nb - 1 || bi % 64 == 0) {
        HIPCHK(hipMemcpyAsync(lt.data(), s->loss_terms, lt.size() * 8, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        double a = 0.0;
        for (double v : lt) a += v;
        lsum = a / bsz;
      }
    }
    if (loss_out) loss_out[ep] = lsum;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  return sur_refresh(s, sub);
}

// Loss gradient of given rows through the training kernels (parity hook).
extern "C" int32_t sce_sur_grad(sce_surrogate* s, int32_t sub, int32_t n, const float* A, const float* B,
                                const double* y, float* grad_out, double* loss) {
  if (!s || sub < 0 || sub >= s->na || !s->subs[sub].params) return sce_fail(SCE_EINVAL, "bad sublattice");
  CHK(sur_scratch(s, n));
  std::vector<float> X((size_t)2 * n * s->fmax);
  std::memcpy(X.data(), A, (size_t)n * s->fmax * 4);
  std::memcpy(X.data() + (size_t)n * s->fmax, B, (size_t)n * s->fmax * 4);
  HIPCHK(hipMemcpy(s->X, X.data(), X.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s->yb, y, (size_t)n * 8, hipMemcpyHostToDevice));
  sur_grad_batch(s, sub, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(grad_out, s->subs[sub].grads, s->subs[sub].np * 4, hipMemcpyDeviceToHost));
  std::vector<double> lt((size_t)n);
  HIPCHK(hipMemcpy(lt.data(), s->loss_terms, lt.size() * 8, hipMemcpyDeviceToHost));
  double a = 0.0;
  for (double v : lt) a += v;
  if (loss) *loss = a / n;
  return SCE_OK;
}

extern "C" int32_t sce_sur_eval(sce_surrogate* s, int32_t sub, int32_t n, const float* A, const float* B,
                                double* out) {
  if (!s || sub < 0 || sub >= s->na || !s->subs[sub].params) return sce_fail(SCE_EINVAL, "bad sublattice");
  Scratch<float> a, b;
  Scratch<double> o;
  CHK(a.up(A, (size_t)n * s->fmax));
  CHK(b.up(B, (size_t)n * s->fmax));
  CHK(o.alloc(n));
  const int nb = (n + 3) / 4;
  const int FP = sur_fpad(s->fmax);
  const size_t lb = (size_t)(TPB / 64) * (2 * FP + 2 * SUR_H) * 4;
  if (s->quant && sur_fpq(s->fmax) > 128)
    return sce_fail(SCE_EINVAL, "int8 stage 1 takes descriptors of at most 128 features (%d here)", s->fmax);
  if (nb && s->quant)
    k_sur_eval<true><<<nb, TPB, lb + da_q_bytes(sur_fpq(s->fmax)), s->stream>>>(s->d_nets, sub, n, s->fmax, FP, a.p,
                                                                                 b.p, o.p);
  else if (nb)
    k_sur_eval<false><<<nb, TPB, lb, s->stream>>>(s->d_nets, sub, n, s->fmax, FP, a.p, b.p, o.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  return o.down(out, n);
}

extern "C" int32_t sce_sur_features(sce_tables* t, const double* config, int32_t n, const int32_t* sites,
                                    const double* e2, float* outA, float* outB) {
  if (!t) return sce_fail(SCE_EINVAL, "NULL tables");
  const DevTables& D = t->dev;
  if (D.lmax < 1) return sce_fail(SCE_EINVAL, "the surrogate descriptor needs lmax >= 1");
  std::vector<int32_t> s0;
  CHK(check_sites(t, n, sites, s0));
  std::vector<double> zr((size_t)D.n_sites * D.nlm);
  CHK(sce_host_zrows(D.lmax, D.n_sites, config, zr.data()));
  Scratch<double> cf, zz, ee;
  Scratch<int32_t> si;
  Scratch<float> a, b;
  CHK(cf.up(config, 3 * (size_t)D.n_sites));
  CHK(zz.up(zr.data(), zr.size()));
  CHK(ee.up(e2, 3 * (size_t)n));
  CHK(si.up(s0.data(), n));
  CHK(a.alloc((size_t)n * D.nfeat_max));
  CHK(b.alloc((size_t)n * D.nfeat_max));
  const int nb = (n + 3) / 4;
  if (nb)
    k_features<false><<<nb, TPB, 4 * 3 * (size_t)std::max(D.nb_max, 1) * 8>>>(D, cf.p, zz.p, nullptr, 0, si.p, ee.p, 0, n,
                                                                              a.p, b.p, nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  CHK(a.down(outA, (size_t)n * D.nfeat_max));
  return b.down(outB, (size_t)n * D.nfeat_max);
}

// ---- the delayed-acceptance sweep ----------------------------------------------
// LDS of a delayed-acceptance block: the four groups, then (int8 inference)
// each group's operands of sur_delta_q.
static size_t da_block_lds(const DevTables& D, bool q) {
  const int ws = D.nb_max * D.nlm + 1 + D.nlm;
  return da_lds_bytes(ws, sur_fpad(D.nfeat_max)) + (q ? da_q_bytes(sur_fpq(D.nfeat_max)) : 0);
}

template <int LMAX>
static void launch_da(const DevTables& D, sce_chain* c, const SurNet* nets, int lo, int n, bool q, hipStream_t s) {
  constexpr int G = TPB / 64, NLM = (LMAX + 1) * (LMAX + 1);
  const int ws = D.nb_max * NLM + 1 + NLM;
  const int FP = sur_fpad(D.nfeat_max);
  const int nb = (n + G - 1) / G;
  if (!nb) return;
  if (q)
    k_da<LMAX, true><<<nb, TPB, da_block_lds(D, true), s>>>(D, c->config, c->zrows, c->prop, c->dE, c->acc, c->att,
                                                             c->st, nets, lo, n, ws, FP);
  else
    k_da<LMAX, false><<<nb, TPB, da_block_lds(D, false), s>>>(D, c->config, c->zrows, c->prop, c->dE, c->acc,
                                                               c->att, c->st, nets, lo, n, ws, FP);
}

template <int LMAX>
static void launch_da_chain(const DevTables& D, sce_chain* c, const SurNet* nets, int lo, int n, PropRange r0,
                            PropRange r1, bool q, hipStream_t s) {
  constexpr int G = TPB / 64, NLM = (LMAX + 1) * (LMAX + 1);
  const int ws = D.nb_max * NLM + 1 + NLM;
  const int FP = sur_fpad(D.nfeat_max);
  const int nwalk = (n + G - 1) / G;
  const int nb = nwalk + prop_blocks(r0.n) + prop_blocks(r1.n);
  if (q)
    k_da_chain<LMAX, true><<<nb, TPB, da_block_lds(D, true), s>>>(D, c->config, c->zrows, c->prop, c->dE, c->acc,
                                                                   c->att, c->st, nets, lo, n, ws, FP, nwalk, r0, r1);
  else
    k_da_chain<LMAX, false><<<nb, TPB, da_block_lds(D, false), s>>>(D, c->config, c->zrows, c->prop, c->dE, c->acc,
                                                                     c->att, c->st, nets, lo, n, ws, FP, nwalk, r0, r1);
}

// One delayed-acceptance sweep, chained within a call as enqueue_metro chains
// plain sweeps (the plan of plan_chain): the next sweep's proposals ride in the
// walk launches; the previous sweep's final (k_final<MODE_DA>, pending >= 0)
// is its own launch before the walks, so no launch lags the sweep count.
static int32_t enqueue_da(sce_chain* c, sce_surrogate* s, int meas, bool need_prop = true, bool next_prop = false,
                          int pending = -1, bool carry_last = false) {
  const DevTables& D = c->t->dev;
  const sce_tables* t = c->t;
  if (da_block_lds(D, s->quant) > 65536)
    return sce_fail(SCE_EINVAL, "delayed-acceptance stencil too large for one block");
  if (s->quant && sur_fpq(D.nfeat_max) > 128)
    return sce_fail(SCE_EINVAL, "int8 stage 1 takes descriptors of at most 128 features (%d here); use FP32",
                    D.nfeat_max);
  if (c->sweeps_done >= 2147483647LL - 1) return sce_fail(SCE_ERANGE_COUNTER, "sweep counter exhausted");
  hipStream_t st = c->stream;
  if (pending >= 0) {
    const FinalArgs f = final_args(c, pending);
    k_final<MODE_DA><<<1, 256, 0, st>>>(f.part, f.n, f.ipart, f.n_sites, f.meas, c->st, f.out, f.n_att, f.n_div,
                                        f.scratch, f.bin);
  }
  ChainPlan P;
  CHK(plan_chain(t->n_colors, t->color_ptr.data(), t->fin_col, next_prop, -1, carry_last, true, &P));
  const PropRange none{0, 0, 0};
  for (const PropRange& r : P.before)
    SCE_DISPATCH_LMAX(D.lmax, launch_propose_range, D, c->config, c->prop, c->st, r, st);
  if (need_prop) SCE_DISPATCH_LMAX(D.lmax, launch_propose, D, c->config, c->prop, c->st, st);
  for (int k = 0; k < P.K; ++k) {
    const int lo = t->color_ptr[P.cols[k]], n = t->color_ptr[P.cols[k] + 1] - lo;
    if (!P.at[k].empty()) {
      SCE_DISPATCH_LMAX(D.lmax, launch_da_chain, D, c, s->d_nets, lo, n, P.at[k][0],
                        P.at[k].size() > 1 ? P.at[k][1] : none, s->quant, st);
    } else {
      SCE_DISPATCH_LMAX(D.lmax, launch_da, D, c, s->d_nets, lo, n, s->quant, st);
    }
  }
  if (next_prop) {
    SCE_DISPATCH_LMAX(D.lmax, launch_reduce_propose_da, c, meas, st, true, P.at[P.K].empty() ? none : P.at[P.K][0]);
  } else {
    launch_reduce<MODE_DA>(c->dE, D.n_active, c->acc, c->att, c->partial, c->ip64, c->config, D.active, D.n_sites,
                           meas, c->st, c->obs, D.n_active, (double)D.n_active, st, c->bin);
  }
  for (const PropRange& r : P.after)
    SCE_DISPATCH_LMAX(D.lmax, launch_propose_range, D, c->config, c->prop, c->st, r, st);
  c->sweeps_done += 1;
  return SCE_OK;
}

// Delayed-acceptance sweeps (surrogate stage + exact FP64 stage) of chain c.
// Same renormalisation / measurement semantics as sce_run_sweeps.
extern "C" int32_t sce_da_run_sweeps(sce_chain* c, sce_surrogate* s, double beta, int64_t nsweeps,
                                     int64_t renorm_interval, int64_t measure_interval, double* obs,
                                     int64_t obs_cap, int64_t* n_obs) {
  CHK(chain_check(c));
  if (!s || s->t != c->t) return sce_fail(SCE_EINVAL, "surrogate and chain on different tables");
  if (nsweeps < 0) return sce_fail(SCE_EINVAL, "nsweeps must be >= 0");
  const int64_t want = measure_interval > 0 ? nsweeps / measure_interval : 0;
  if (want > 0 && (!obs || obs_cap < want)) return sce_fail(SCE_EDIM, "obs buffer too small");
  if (c->sweeps_done + nsweeps >= 2147483647LL - 1) return sce_fail(SCE_ERANGE_COUNTER, "sweep counter exhausted");
  if (want > 0 && (size_t)want > c->obs_cap) {
    const size_t cap = obs_rows_for((size_t)want, c->obs_cap);
    double* p;
    CHK(dalloc(c->mem, &p, 4 * cap));
    c->drop_graphs();
    c->mem.release(c->obs);
    c->obs = p;
    c->obs_cap = cap;
  }
  CHK(set_beta(c, beta));
  {
    ChainDev h;
    CHK(read_state(c, &h));
    h.n_obs = 0;
    HIPCHK(hipMemcpy(c->st, &h, sizeof h, hipMemcpyHostToDevice));
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  // Chained within the call as plain sweeps are (enqueue_da); with graph
  // replay each variant [fused measurement + 2 need_prop + 4 next_prop + 8
  // measurement flag of the pending final] is captured on first use.
  GraphSlot ge[16];
  auto drop = [&] {
    (void)hipStreamSynchronize(c->stream);
    for (auto& x : ge) x.reset();
  };
  int32_t rc = SCE_OK;
  bool have_prop = false;
  int pending = -1;
  for (int64_t i = 1; i <= nsweeps && rc == SCE_OK; ++i) {
    const bool renorm = renorm_interval > 0 && i % renorm_interval == 0;
    const bool meas = measure_interval > 0 && i % measure_interval == 0;
    const bool fused = meas && !renorm;
    const bool next_prop = !renorm && i < nsweeps;
    if (c->use_graph) {
      const int k = (fused ? 1 : 0) + (have_prop ? 0 : 2) + (next_prop ? 4 : 0) + (pending == 1 ? 8 : 0);
      if (!ge[k]) {
        const long long before = c->sweeps_done;
        const bool chained = (k & 2) == 0;
        rc = capture_graph(c->stream, ge[k], [&] {
          return enqueue_da(c, s, k & 1, !chained, (k & 4) != 0, chained ? (k >> 3) & 1 : -1, chained);
        });
        c->sweeps_done = before;
        if (rc != SCE_OK) break;
      }
      if (hipGraphLaunch(ge[k].x, c->stream) != hipSuccess) rc = sce_fail(SCE_EHIP, "hipGraphLaunch failed");
      c->sweeps_done += 1;
    } else {
      rc = enqueue_da(c, s, fused ? 1 : 0, !have_prop, next_prop, pending, have_prop);
    }
    have_prop = next_prop;
    pending = next_prop ? (fused ? 1 : 0) : -1;
    if (renorm) enqueue_renorm(c);
    if (meas && !fused) enqueue_measure(c);
  }
  drop();  // drains the stream first
  CHK(rc);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  if (want > 0) HIPCHK(hipMemcpy(obs, c->obs, 4 * (size_t)want * 8, hipMemcpyDeviceToHost));
  if (n_obs) *n_obs = want;
  return SCE_OK;
}

extern "C" int32_t sce_test_mfma(int64_t n, const float* a4, const float* b4, const float* c, float* out) {
  CHK(need_device());
  if (n < 0 || n > (1 << 20)) return sce_fail(SCE_EINVAL, "n out of range");
  Scratch<float> a, b, cc, o;
  CHK(a.up(a4, 4 * (size_t)n));
  CHK(b.up(b4, 4 * (size_t)n));
  CHK(cc.up(c, (size_t)n));
  CHK(o.alloc((size_t)n));
  if (n) k_test_mfma<<<(unsigned)n, 64>>>(a.p, b.p, cc.p, o.p, (int)n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return o.down(out, (size_t)n);
}
