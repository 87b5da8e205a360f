/*
 * sce_oracle.c — CPU restatement of SCEMonteCarlo.jl's device Metropolis path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the timed CPU
 * baseline.  Compiled with -ffp-contract=off and without fast-math so that every
 * floating-point operation is the IEEE operation written here.  Each function
 * cites the reference file:line it restates.
 */
#include "sce_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * Philox4x32-10 — src/gpu/philox.jl:13-16 (constants), :26-35 (round),
 * :43-50 (10 rounds, Weyl key bump), :55-60 (counter layout), :78-81 (uniform),
 * :100-105 (Box-Muller).
 * ====================================================================== */
#define PH_M0 0xd2511f53u
#define PH_M1 0xcd9e8d57u
#define PH_W0 0x9e3779b9u
#define PH_W1 0xbb67ae85u

static void ph_round(uint32_t c[4], const uint32_t k[2]) {
  uint64_t p0 = (uint64_t)PH_M0 * (uint64_t)c[0];
  uint64_t p1 = (uint64_t)PH_M1 * (uint64_t)c[2];
  uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  uint32_t n0 = hi1 ^ c[1] ^ k[0];
  uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void ora_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 9; ++r) {
    ph_round(c, k);
    k[0] += PH_W0;
    k[1] += PH_W1;
  }
  ph_round(c, k);
  memcpy(out, c, sizeof c);
}

/* ctr = (site, sweep, slot, tag) with site/sweep reinterpreted Int32 -> UInt32;
 * key = (lo32(seed), hi32(seed)).  The reference reserves ctr[4] (tag) = 0 for
 * its Metropolis streams (philox.jl:52-54). */
void ora_philox_block(uint64_t seed, int32_t site, int32_t sweep, uint32_t slot,
                      uint32_t tag, uint32_t out[4]) {
  uint32_t ctr[4] = {(uint32_t)site, (uint32_t)sweep, slot, tag};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  ora_philox4x32(ctr, key, out);
}

double ora_uniform(uint32_t hi, uint32_t lo) {
  uint64_t w = ((uint64_t)hi << 32) | (uint64_t)lo;
  return ((double)(w >> 12) + 0.5) * 0x1p-52;
}

/* ======================================================================
 * Deterministic elementary functions.  The reference's Box-Muller, rotation
 * and accept test call the backend libm (philox.jl:96-98, gpu_sweep.jl:136),
 * which makes host and device trajectories differ in ULPs.  This library instead
 * defines exp/log/sincos by named published methods — Cody & Waite argument
 * reduction (Software Manual for the Elementary Functions, 1980) followed by
 * the Taylor / atanh series truncated below 1e-17 relative — evaluated with
 * plain IEEE +,-,*,/ in a fixed order, so that any IEEE host and the GPU give
 * identical bits.  Coefficients are 1/k! and 1/(2k+1) rounded to double.
 * ====================================================================== */
static const double LN2_HI = 0x1.62e42fee00000p-1;
static const double LN2_LO = 0x1.a39ef35793c76p-33;
static const double INV_LN2 = 0x1.71547652b82fep+0;
static const double PIO2_1 = 0x1.921fb54400000p+0;
static const double PIO2_2 = 0x1.0b4611a600000p-34;
static const double PIO2_3 = 0x1.3198a2e037073p-69;
static const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
static const double SQRT2 = 0x1.6a09e667f3bcdp+0;

static const double EXP_C[14] = {
    0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000000p-1,
    0x1.5555555555555p-3, 0x1.5555555555555p-5, 0x1.1111111111111p-7,
    0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16,
    0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26,
    0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33};
static const double LOG_C[12] = {
    0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3,
    0x1.c71c71c71c71cp-4, 0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4,
    0x1.1111111111111p-4, 0x1.e1e1e1e1e1e1ep-5, 0x1.af286bca1af28p-5,
    0x1.8618618618618p-5, 0x1.642c8590b2164p-5, 0x1.47ae147ae147bp-5};
static const double SIN_C[9] = {
    -0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13,
    0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33,
    -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49, -0x1.2f49b46814157p-57};
static const double COS_C[10] = {
    -0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10,
    0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29,
    -0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-45, -0x1.6827863b97d97p-53,
    0x1.e542ba4020225p-62};

static double bits2d(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t d2bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }
static double pow2i(int k) { return bits2d((uint64_t)(k + 1023) << 52); } /* -1022..1023 */

double ora_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return INFINITY;
  if (x < -745.2) return 0.0;
  double n = rint(x * INV_LN2);
  double r = (x - n * LN2_HI) - n * LN2_LO;
  double p = EXP_C[13];
  for (int k = 12; k >= 0; --k) p = p * r + EXP_C[k];
  int k = (int)n;
  if (k > 1023) return p * pow2i(1023) * pow2i(k - 1023);
  if (k < -1022) return p * pow2i(k + 1000) * pow2i(-1000);
  return p * pow2i(k);
}

double ora_log(double x) {
  if (x != x || x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (x == INFINITY) return x;
  int eadj = 0;
  if (x < 0x1p-1022) { x = x * 0x1p54; eadj = -54; }
  uint64_t b = d2bits(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023 + eadj;
  double m = bits2d((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > SQRT2) { m = m * 0.5; e += 1; }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double t = LOG_C[11];
  for (int k = 10; k >= 0; --k) t = t * z + LOG_C[k];
  double R = t * z;
  double s2 = s + s;
  double lm = s2 + s2 * R;
  double de = (double)e;
  return de * LN2_HI + (de * LN2_LO + lm);
}

void ora_sincos(double x, double* so, double* co) {
  double n = rint(x * TWO_OVER_PI);
  double r = ((x - n * PIO2_1) - n * PIO2_2) - n * PIO2_3;
  double z = r * r;
  double ts = SIN_C[8];
  for (int k = 7; k >= 0; --k) ts = ts * z + SIN_C[k];
  double sn = r + (r * z) * ts;
  double tc = COS_C[9];
  for (int k = 8; k >= 0; --k) tc = tc * z + COS_C[k];
  double cs = 1.0 + z * tc;
  int q = (int)((long long)n & 3);
  switch (q) {
    case 0: *so = sn; *co = cs; break;
    case 1: *so = cs; *co = -sn; break;
    case 2: *so = -sn; *co = -cs; break;
    default: *so = -cs; *co = sn; break;
  }
}

/* Box-Muller from one block — philox.jl:100-105, with the project log/sincos. */
void ora_normal2(const uint32_t blk[4], double out[2]) {
  double u1 = ora_uniform(blk[0], blk[1]);
  double u2 = ora_uniform(blk[2], blk[3]);
  double r = sqrt(-2.0 * ora_log(u1));
  double a = 6.283185307179586 * u2;
  double s, c;
  ora_sincos(a, &s, &c);
  out[0] = r * c;
  out[1] = r * s;
}

/* ======================================================================
 * Tesseral row — src/gpu/zlm_device.jl:151-263 (replica of Harmonics.Zlm_unsafe
 * -> LegendrePolynomials.dnPl v0.4, operation order preserved).
 * ====================================================================== */
static double zlm_plm_norm(int l, int m) { /* zlm_device.jl:154-160 */
  double acc = (double)(2 * l + 1) / (4.0 * 3.141592653589793);
  for (int i = l - m + 1; i <= l + m; ++i) acc /= (double)i;
  return sqrt(acc);
}

typedef struct { double re, im; } cplx;
static cplx cmul(cplx a, cplx b) {
  cplx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}
static int tz(unsigned p) { int t = 0; while (!(p & 1u)) { p >>= 1; ++t; } return t; }

static cplx zlm_cpow(cplx x, int p) { /* zlm_device.jl:168-191 */
  if (p == 0) { cplx one = {1.0, 0.0}; return one; }
  cplx xsq = cmul(x, x);
  if (p == 1) return x;
  if (p == 2) return xsq;
  int t = tz((unsigned)p) + 1;
  p >>= t;
  if ((t -= 1) > 0) x = xsq;
  while ((t -= 1) > 0) x = cmul(x, x);
  cplx y = x;
  while (p > 0) {
    t = tz((unsigned)p) + 1;
    p >>= t;
    while ((t -= 1) >= 0) x = cmul(x, x);
    y = cmul(y, x);
  }
  return y;
}

static double zlm_dnpl(double x, int l, int n, double* cache) { /* :200-232 */
  if (n == l) {
    double p = 1.0;
    for (int i = 2 * l - 1; i >= 1; i -= 2) p *= (double)i;
    cache[0] = p;
  } else {
    double Pl = 1.0, Plm1 = 0.0;
    cache[0] = Pl;
    for (int ll = 1; ll <= l - n; ++ll) {
      double nPl = ((double)(2 * ll - 1) * x * Pl - (double)(ll - 1) * Plm1) / (double)ll;
      Plm1 = Pl;
      Pl = nPl;
      cache[ll] = Pl;
    }
    for (int ni = 1; ni <= n; ++ni) {
      double pnn = (double)(2 * ni - 1) * cache[0];
      cache[0] = pnn;
      int ll = ni + 1;
      cache[1] = ((double)(2 * ll - 1) * (x * pnn + (double)ni * cache[1])) / (double)ll;
      int top = (l < l - n + ni) ? l : l - n + ni;
      for (int li = ni + 2; li <= top; ++li) {
        /* cache[li-ni+1] (1-based) = cache[li-ni] (0-based) */
        cache[li - ni] = ((double)(2 * li - 1) * (x * cache[li - ni - 1] +
                                                  (double)ni * cache[li - ni]) -
                          (double)(li - 1) * cache[li - ni - 2]) / (double)li;
      }
    }
  }
  return cache[l - n];
}

static double zlm_one(int l, int m, const double u[3], double* cache) { /* :235-243 */
  int n = m < 0 ? -m : m;
  double par = (n & 1) ? -1.0 : 1.0;
  double plm = par * zlm_plm_norm(l, n) * zlm_dnpl(u[2], l, n, cache);
  if (n == 0) return plm;
  double c = par * sqrt(2.0) * plm;
  cplx xy = {u[0], u[1]};
  cplx zp = zlm_cpow(xy, n);
  return m > 0 ? c * zp.re : c * zp.im;
}

void ora_zlm_row(const double u[3], int lmax, double* z) { /* :254-263 */
  double cache[8];
  int i = 0;
  for (int l = 0; l <= lmax; ++l)
    for (int m = -l; m <= l; ++m) z[i++] = zlm_one(l, m, u, cache);
}

/* The tangent-projected gradient of one tesseral harmonic, grad Z_lm(u)
 * (gpu/grad_device.jl:30-66, the device twin of Harmonics.grad_Zlm_unsafe):
 * plm and dplm are the order-n and order-(n+1) derivatives of P_l at u_z with
 * the parity and normalisation factors (the order-(n+1) derivative is a +0.0
 * literal when n + 1 > l), then the Cartesian derivative of
 * c P_l^(n)(z) Re/Im (x + iy)^n with the radial part u (u . dZ) / |u|^2
 * removed.  Only + - * / and a correctly rounded sqrt. */
static void grad_zlm_one(int l, int m, const double u[3], double* cache, double g[3]) {
  const double x = u[0], y = u[1], z = u[2];
  const int n = m < 0 ? -m : m;
  const double par = (n & 1) ? -1.0 : 1.0;
  const double plm = par * zlm_plm_norm(l, n) * zlm_dnpl(z, l, n, cache);
  const double d1 = l < n + 1 ? 0.0 : zlm_dnpl(z, l, n + 1, cache);
  const double dplm = par * zlm_plm_norm(l, n) * d1;
  const double r2 = x * x + y * y + z * z;
  if (m == 0) {
    const double zz = z * dplm / r2;
    g[0] = -x * zz;
    g[1] = -y * zz;
    g[2] = dplm - z * zz;
    return;
  }
  const double c = par * sqrt(2.0);
  cplx xy = {x, y};
  const cplx pn = zlm_cpow(xy, n), pn1 = zlm_cpow(xy, n - 1);
  double dx, dy, dz;
  if (m > 0) {
    dx = c * (double)n * plm * pn1.re;
    dy = -c * (double)n * plm * pn1.im;
    dz = c * dplm * pn.re;
  } else {
    dx = c * (double)n * plm * pn1.im;
    dy = c * (double)n * plm * pn1.re;
    dz = c * dplm * pn.im;
  }
  const double zz = (x * dx + y * dy + z * dz) / r2;
  g[0] = dx - x * zz;
  g[1] = dy - y * zz;
  g[2] = dz - z * zz;
}

/* grad Z rows of u in lm order, component fastest: grow[3k + d]. */
void ora_grad_zlm_row(const double u[3], int lmax, double* grow) {
  double cache[8];
  int i = 0;
  for (int l = 0; l <= lmax; ++l)
    for (int m = -l; m <= l; ++m, ++i) grad_zlm_one(l, m, u, cache, grow + 3 * i);
}

/* ======================================================================
 * Tiled Hamiltonian — src/hamiltonian.jl:99-181 (programs), :267-393 (ctor),
 * :437-448 (Zeeman templates), :456-499 (greedy colouring), :520-524 (site id).
 * ====================================================================== */
typedef struct { int32_t* a; int64_t n, cap; } vi32;
typedef struct { double* a; int64_t n, cap; } vf64;
typedef struct { int8_t* a; int64_t n, cap; } vi8;
#define VPUSH(v, x)                                                        \
  do {                                                                     \
    if ((v).n == (v).cap) {                                                \
      (v).cap = (v).cap ? 2 * (v).cap : 64;                                \
      (v).a = realloc((v).a, (size_t)(v).cap * sizeof(*(v).a));            \
    }                                                                      \
    (v).a[(v).n++] = (x);                                                  \
  } while (0)

typedef struct {
  int body;
  double coef;
  int32_t atoms[8];
  int32_t sh[8][3];
  int32_t ls[8];
  const double* folded;
  double* owned;
} oterm;

struct ora_tiled {
  int n_cell_atoms, dims[3], n_sites, lmax, nlm, n_active, n_colors, n_terms,
      n_fitted, n_inst;
  vi32 inst_term, inst_ptr, inst_sites, site_ptr, site_inst, site_prog, sprog_ptr,
      sent_tgt, sfac_ptr, sfac_row, site_col, site_col2, pent_row, pent_row2,
      eprog_ptr, efac_ptr, efac_row, color_ptr, color_sites;
  vi8 site_slot, sfac_slot, site_has_l1, site_active;
  vf64 sent_w, eent_w, term_coef;
  oterm* terms;
};

static int lm_index(int l, int m) { return l * l + l + m + 1; }

static int tiled_err(char* err, size_t len, const char* fmt, ...) {
  if (err && len) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, len, fmt, ap);
    va_end(ap);
  }
  return -1;
}

static int imod(int x, int d) { int r = x % d; return r < 0 ? r + d : r; }

static int32_t site_index(int na, const int d[3], int atom, const int c[3]) {
  return (int32_t)(atom + na * (c[0] + d[0] * (c[1] + d[1] * c[2])));
}

/* hamiltonian.jl:99-133 — site programs (one per member slot) then the energy
 * program, CartesianIndices column-major over `folded`. */
static void push_term_programs(ora_tiled* t, const oterm* tm) {
  int D = tm->body;
  int dimv[8], total = 1;
  for (int k = 0; k < D; ++k) { dimv[k] = 2 * tm->ls[k] + 1; total *= dimv[k]; }
  int idx[8];
  for (int v = 0; v < D; ++v) {
    for (int lin = 0; lin < total; ++lin) {
      int r = lin;
      for (int k = 0; k < D; ++k) { idx[k] = r % dimv[k]; r /= dimv[k]; }
      double w = tm->coef * tm->folded[lin];
      if (w == 0.0) continue;
      VPUSH(t->sent_w, w);
      VPUSH(t->sent_tgt, lm_index(tm->ls[v], idx[v] - tm->ls[v]));
      for (int k = 0; k < D; ++k) {
        if (k == v) continue;
        VPUSH(t->sfac_row, lm_index(tm->ls[k], idx[k] - tm->ls[k]));
        VPUSH(t->sfac_slot, (int8_t)(k + 1));
      }
      VPUSH(t->sfac_ptr, (int32_t)(t->sfac_row.n + 1));
      int32_t pr = 0, pr2 = 0;
      if (D == 2) pr = t->sfac_row.a[t->sfac_row.n - 1];
      if (D == 3) { pr = t->sfac_row.a[t->sfac_row.n - 2]; pr2 = t->sfac_row.a[t->sfac_row.n - 1]; }
      VPUSH(t->pent_row, pr);
      VPUSH(t->pent_row2, pr2);
    }
    VPUSH(t->sprog_ptr, (int32_t)(t->sent_w.n + 1));
  }
  for (int lin = 0; lin < total; ++lin) {
    int r = lin;
    for (int k = 0; k < D; ++k) { idx[k] = r % dimv[k]; r /= dimv[k]; }
    double w = tm->folded[lin];
    if (w == 0.0) continue;
    VPUSH(t->eent_w, w);
    for (int k = 0; k < D; ++k) VPUSH(t->efac_row, lm_index(tm->ls[k], idx[k] - tm->ls[k]));
    VPUSH(t->efac_ptr, (int32_t)(t->efac_row.n + 1));
  }
  VPUSH(t->eprog_ptr, (int32_t)(t->eent_w.n + 1));
}

int ora_tiled_build(int na, int n_terms, const int32_t* body, const double* coef,
                    const int32_t* atoms, const int32_t* shifts, const int32_t* ls,
                    const double* folded, const int32_t dims_in[3],
                    const double* magmoms, const double* field, ora_tiled** out,
                    char* err, size_t errlen) {
  *out = NULL;
  if (na < 1) return tiled_err(err, errlen, "n_cell_atoms must be >= 1");
  int d[3] = {dims_in[0], dims_in[1], dims_in[2]};
  if (d[0] < 1 || d[1] < 1 || d[2] < 1) return tiled_err(err, errlen, "dims must be >= 1");
  const double MU_B_EV_T = 9.2740100783e-24 / 1.602176634e-19;
  const double FOURPI = 4.0 * 3.141592653589793;
  const double N1 = sqrt(3.0 / FOURPI);
  int have_field = field && (field[0] != 0.0 || field[1] != 0.0 || field[2] != 0.0);
  if (field && !magmoms) return tiled_err(err, errlen, "field given without magmoms");
  oterm* terms = calloc((size_t)(n_terms + na + 1), sizeof(oterm));
  int nt = 0, lmax = 0;
  int64_t ao = 0, fo = 0;
  for (int k = 0; k < n_terms; ++k) {
    int D = body[k];
    if (D < 1 || D > 8) { free(terms); return tiled_err(err, errlen, "term %d: body %d unsupported", k + 1, D); }
    int total = 1;
    for (int i = 0; i < D; ++i) total *= 2 * ls[ao + i] + 1;
    if (coef[k] != 0.0) { /* hamiltonian.jl:280 drops coef == 0 */
      oterm* tm = &terms[nt];
      tm->body = D;
      for (int i = 0; i < D; ++i) {
        tm->atoms[i] = atoms[ao + i];
        tm->ls[i] = ls[ao + i];
        for (int x = 0; x < 3; ++x) tm->sh[i][x] = shifts[3 * (ao + i) + x];
        if (tm->atoms[i] < 1 || tm->atoms[i] > na) { free(terms); return tiled_err(err, errlen, "term %d: atom outside 1:%d", k + 1, na); }
        if (tm->ls[i] < 0) { free(terms); return tiled_err(err, errlen, "term %d: negative l", k + 1); }
        if (tm->ls[i] > lmax) lmax = tm->ls[i];
      }
      if (tm->sh[0][0] || tm->sh[0][1] || tm->sh[0][2]) { free(terms); return tiled_err(err, errlen, "term %d: shifts[1] != 0", k + 1); }
      for (int i = 0; i < D; ++i)
        for (int j = i + 1; j < D; ++j) {
          int same = tm->atoms[i] == tm->atoms[j];
          for (int x = 0; x < 3; ++x) same &= imod(tm->sh[i][x], d[x]) == imod(tm->sh[j][x], d[x]);
          if (same) { free(terms); return tiled_err(err, errlen, "term %d folds two member sites onto one supercell site; enlarge dims", k + 1); }
        }
      tm->coef = coef[k] * pow(FOURPI, (double)D / 2.0); /* hamiltonian.jl:316 */
      tm->folded = folded + fo;
      ++nt;
    }
    ao += D;
    fo += total;
  }
  int n_fitted = nt;
  if (magmoms && have_field) { /* hamiltonian.jl:437-448 */
    for (int a = 0; a < na; ++a) {
      double m = magmoms[a];
      if (m == 0.0) continue;
      oterm* tm = &terms[nt++];
      tm->body = 1;
      tm->coef = 1.0;
      tm->atoms[0] = a + 1;
      tm->ls[0] = 1;
      tm->owned = malloc(3 * sizeof(double));
      double s = -(MU_B_EV_T * m / N1);
      tm->owned[0] = s * field[1];
      tm->owned[1] = s * field[2];
      tm->owned[2] = s * field[0];
      tm->folded = tm->owned;
    }
    if (nt > n_fitted && lmax < 1) lmax = 1;
  }
  if (nt == 0) { free(terms); return tiled_err(err, errlen, "the term list is empty"); }

  ora_tiled* t = calloc(1, sizeof(ora_tiled));
  t->terms = terms;
  t->n_cell_atoms = na;
  memcpy(t->dims, d, sizeof d);
  int ncells = d[0] * d[1] * d[2];
  t->n_sites = na * ncells;
  t->lmax = lmax;
  t->nlm = (lmax + 1) * (lmax + 1);
  t->n_terms = nt;
  t->n_fitted = n_fitted;
  t->n_inst = nt * ncells;

  /* instances: cell-outer (c3, c2, c1 with c1 fastest), term-inner — :333-353 */
  VPUSH(t->inst_ptr, 1);
  for (int c3 = 0; c3 < d[2]; ++c3)
    for (int c2 = 0; c2 < d[1]; ++c2)
      for (int c1 = 0; c1 < d[0]; ++c1)
        for (int k = 0; k < nt; ++k) {
          VPUSH(t->inst_term, k + 1);
          for (int i = 0; i < terms[k].body; ++i) {
            int cw[3] = {imod(c1 + terms[k].sh[i][0], d[0]),
                         imod(c2 + terms[k].sh[i][1], d[1]),
                         imod(c3 + terms[k].sh[i][2], d[2])};
            VPUSH(t->inst_sites, site_index(na, d, terms[k].atoms[i], cw));
          }
          VPUSH(t->inst_ptr, (int32_t)(t->inst_sites.n + 1));
        }
  /* site CSR: count, prefix, fill — :355-375 */
  int N = t->n_sites;
  int32_t* counts = calloc((size_t)N + 1, sizeof(int32_t));
  for (int64_t q = 0; q < t->inst_sites.n; ++q) counts[t->inst_sites.a[q] - 1]++;
  VPUSH(t->site_ptr, 1);
  for (int s = 0; s < N; ++s) VPUSH(t->site_ptr, t->site_ptr.a[s] + counts[s]);
  int64_t tot = t->inst_sites.n;
  t->site_inst.a = malloc((size_t)tot * 4); t->site_inst.n = t->site_inst.cap = tot;
  t->site_slot.a = malloc((size_t)tot); t->site_slot.n = t->site_slot.cap = tot;
  int32_t* cursor = malloc((size_t)N * 4);
  for (int s = 0; s < N; ++s) cursor[s] = t->site_ptr.a[s];
  for (int i = 0; i < t->n_inst; ++i) {
    int slot = 0;
    for (int q = t->inst_ptr.a[i]; q < t->inst_ptr.a[i + 1]; ++q) {
      ++slot;
      int s = t->inst_sites.a[q - 1];
      t->site_inst.a[cursor[s - 1] - 1] = i + 1;
      t->site_slot.a[cursor[s - 1] - 1] = (int8_t)slot;
      cursor[s - 1]++;
    }
  }
  /* has_l1, active — :377-383 */
  t->site_has_l1.a = calloc((size_t)N, 1); t->site_has_l1.n = N;
  t->site_active.a = calloc((size_t)N, 1); t->site_active.n = N;
  t->n_active = 0;
  for (int s = 0; s < N; ++s) {
    for (int j = t->site_ptr.a[s]; j < t->site_ptr.a[s + 1]; ++j) {
      const oterm* tm = &terms[t->inst_term.a[t->site_inst.a[j - 1] - 1] - 1];
      if (tm->ls[t->site_slot.a[j - 1] - 1] == 1) t->site_has_l1.a[s] = 1;
    }
    t->site_active.a[s] = t->site_ptr.a[s + 1] > t->site_ptr.a[s];
    t->n_active += t->site_active.a[s];
  }
  /* programs — :138-181 */
  VPUSH(t->sprog_ptr, 1); VPUSH(t->sfac_ptr, 1); VPUSH(t->eprog_ptr, 1); VPUSH(t->efac_ptr, 1);
  int np = 0;
  int* pbase = malloc((size_t)nt * sizeof(int));
  vi8 pslot1 = {0}, pslot2 = {0};
  for (int k = 0; k < nt; ++k) {
    pbase[k] = np;
    int D = terms[k].body;
    np += D;
    for (int v = 1; v <= D; ++v) {
      VPUSH(pslot1, (int8_t)(D == 2 ? 3 - v : D == 3 ? (v == 1 ? 2 : 1) : 0));
      VPUSH(pslot2, (int8_t)(D == 3 ? (v == 3 ? 2 : 3) : 0));
    }
    push_term_programs(t, &terms[k]);
    VPUSH(t->term_coef, terms[k].coef);
  }
  t->site_prog.a = malloc((size_t)tot * 4); t->site_prog.n = tot;
  t->site_col.a = malloc((size_t)tot * 4); t->site_col.n = tot;
  t->site_col2.a = malloc((size_t)tot * 4); t->site_col2.n = tot;
  for (int64_t j = 0; j < tot; ++j) {
    int inst = t->site_inst.a[j];
    int pid = pbase[t->inst_term.a[inst - 1] - 1] + t->site_slot.a[j];
    t->site_prog.a[j] = pid;
    int off = t->inst_ptr.a[inst - 1] - 1;
    int sl1 = pslot1.a[pid - 1], sl2 = pslot2.a[pid - 1];
    t->site_col.a[j] = sl1 == 0 ? 0 : sl2 == 0 ? t->inst_sites.a[off + sl1 - 1]
                                               : -t->inst_sites.a[off + sl1 - 1];
    t->site_col2.a[j] = sl2 == 0 ? 0 : t->inst_sites.a[off + sl2 - 1];
  }
  free(pbase); free(pslot1.a); free(pslot2.a);

  /* greedy colouring — :456-499 */
  int32_t* colors = calloc((size_t)N, 4);
  vi32 stamp = {0};
  int ncol = 0;
  for (int s = 1; s <= N; ++s) {
    if (!t->site_active.a[s - 1]) continue;
    for (int j = t->site_ptr.a[s - 1]; j < t->site_ptr.a[s]; ++j) {
      int i = t->site_inst.a[j - 1];
      for (int q = t->inst_ptr.a[i - 1]; q < t->inst_ptr.a[i]; ++q) {
        int c = colors[t->inst_sites.a[q - 1] - 1];
        if (c > 0) stamp.a[c - 1] = s;
      }
    }
    int c = 1;
    while (c <= ncol && stamp.a[c - 1] == s) ++c;
    if (c > ncol) { ncol = c; VPUSH(stamp, 0); }
    colors[s - 1] = c;
  }
  free(stamp.a);
  memset(counts, 0, (size_t)N * 4);
  for (int s = 0; s < N; ++s) if (colors[s] > 0) counts[colors[s] - 1]++;
  VPUSH(t->color_ptr, 1);
  for (int c = 0; c < ncol; ++c) VPUSH(t->color_ptr, t->color_ptr.a[c] + counts[c]);
  t->color_sites.a = malloc((size_t)(t->n_active + 1) * 4);
  t->color_sites.n = t->n_active;
  for (int c = 0; c < ncol; ++c) cursor[c] = t->color_ptr.a[c];
  for (int s = 1; s <= N; ++s) {
    int c = colors[s - 1];
    if (c <= 0) continue;
    t->color_sites.a[cursor[c - 1] - 1] = s;
    cursor[c - 1]++;
  }
  t->n_colors = ncol;
  free(colors); free(counts); free(cursor);
  *out = t;
  return 0;
}

void ora_tiled_free(ora_tiled* t) {
  if (!t) return;
  vi32* v32[] = {&t->inst_term, &t->inst_ptr, &t->inst_sites, &t->site_ptr, &t->site_inst,
                 &t->site_prog, &t->sprog_ptr, &t->sent_tgt, &t->sfac_ptr, &t->sfac_row,
                 &t->site_col, &t->site_col2, &t->pent_row, &t->pent_row2, &t->eprog_ptr,
                 &t->efac_ptr, &t->efac_row, &t->color_ptr, &t->color_sites};
  for (size_t i = 0; i < sizeof v32 / sizeof *v32; ++i) free(v32[i]->a);
  free(t->site_slot.a); free(t->sfac_slot.a); free(t->site_has_l1.a); free(t->site_active.a);
  free(t->sent_w.a); free(t->eent_w.a); free(t->term_coef.a);
  for (int k = 0; k < t->n_terms; ++k) free(t->terms[k].owned);
  free(t->terms);
  free(t);
}

void ora_tiled_scalars(const ora_tiled* t, int64_t o[12]) {
  o[0] = t->n_cell_atoms; o[1] = t->dims[0]; o[2] = t->dims[1]; o[3] = t->dims[2];
  o[4] = t->n_sites; o[5] = t->lmax; o[6] = t->nlm; o[7] = t->n_active;
  o[8] = t->n_colors; o[9] = t->n_terms; o[10] = t->n_fitted; o[11] = t->n_inst;
}

const void* ora_tiled_array(const ora_tiled* t, int f, int64_t* len) {
#define RET(v) do { *len = (v).n; return (v).a; } while (0)
  switch (f) {
    case ORA_INST_TERM: RET(t->inst_term);
    case ORA_INST_PTR: RET(t->inst_ptr);
    case ORA_INST_SITES: RET(t->inst_sites);
    case ORA_SITE_PTR: RET(t->site_ptr);
    case ORA_SITE_INST: RET(t->site_inst);
    case ORA_SITE_SLOT: RET(t->site_slot);
    case ORA_SITE_HAS_L1: RET(t->site_has_l1);
    case ORA_SITE_ACTIVE: RET(t->site_active);
    case ORA_SITE_PROG: RET(t->site_prog);
    case ORA_SPROG_PTR: RET(t->sprog_ptr);
    case ORA_SENT_W: RET(t->sent_w);
    case ORA_SENT_TGT: RET(t->sent_tgt);
    case ORA_SFAC_PTR: RET(t->sfac_ptr);
    case ORA_SFAC_ROW: RET(t->sfac_row);
    case ORA_SFAC_SLOT: RET(t->sfac_slot);
    case ORA_SITE_COL: RET(t->site_col);
    case ORA_SITE_COL2: RET(t->site_col2);
    case ORA_PENT_ROW: RET(t->pent_row);
    case ORA_PENT_ROW2: RET(t->pent_row2);
    case ORA_EPROG_PTR: RET(t->eprog_ptr);
    case ORA_EENT_W: RET(t->eent_w);
    case ORA_EFAC_PTR: RET(t->efac_ptr);
    case ORA_EFAC_ROW: RET(t->efac_row);
    case ORA_TERM_COEF: RET(t->term_coef);
    case ORA_COLOR_PTR: RET(t->color_ptr);
    case ORA_COLOR_SITES: RET(t->color_sites);
  }
#undef RET
  *len = 0;
  return NULL;
}

/* ======================================================================
 * Energy contract — src/energy.jl:41-50 (_zrows), :55-75 (_total_energy),
 * :100-139 (site_coeffs!), :148-157 (delta_energy).
 * ====================================================================== */
#define ZR(z, nlm, row, col) (z)[(size_t)((col) - 1) * (size_t)(nlm) + (size_t)((row) - 1)]

void ora_zrows(const ora_tiled* t, const double* config, double* zrows) {
  for (int s = 0; s < t->n_sites; ++s)
    ora_zlm_row(config + 3 * (size_t)s, t->lmax, zrows + (size_t)s * t->nlm);
}

static double instance_energy(const ora_tiled* t, int i, const double* zrows) {
  int k = t->inst_term.a[i];
  int off = t->inst_ptr.a[i] - 1;
  double Ei = 0.0;
  for (int e = t->eprog_ptr.a[k - 1]; e < t->eprog_ptr.a[k]; ++e) {
    double p = 1.0;
    int f0 = t->efac_ptr.a[e - 1] - 1;
    int nf = t->efac_ptr.a[e] - 1 - f0;
    for (int m = 1; m <= nf; ++m)
      p *= ZR(zrows, t->nlm, t->efac_row.a[f0 + m - 1], t->inst_sites.a[off + m - 1]);
    Ei += t->eent_w.a[e - 1] * p;
  }
  return t->term_coef.a[k - 1] * Ei;
}

double ora_total_energy(const ora_tiled* t, const double* zrows) {
  double E = 0.0;
  for (int i = 0; i < t->n_inst; ++i) E += instance_energy(t, i, zrows);
  return E;
}

/* Fixed-shape pairwise tree over 512-element chunks (zero padded), recursing
 * over the chunk sums: the device's deterministic reduction. */
#define TREE_B 512
static double tree_chunk(const double* x, int64_t n) {
  double v[TREE_B];
  for (int i = 0; i < TREE_B; ++i) v[i] = i < n ? x[i] : 0.0;
  for (int off = TREE_B / 2; off >= 1; off >>= 1)
    for (int i = 0; i < off; ++i) v[i] = v[i] + v[i + off];
  return v[0];
}
double ora_reduce_tree(int64_t n, const double* x) {
  if (n <= TREE_B) return tree_chunk(x, n);
  int64_t nb = (n + TREE_B - 1) / TREE_B;
  double* p = malloc((size_t)nb * sizeof(double));
  for (int64_t b = 0; b < nb; ++b) {
    int64_t m = n - b * TREE_B;
    p[b] = tree_chunk(x + b * TREE_B, m < TREE_B ? m : TREE_B);
  }
  double r = ora_reduce_tree(nb, p);
  free(p);
  return r;
}

double ora_total_energy_tree(const ora_tiled* t, const double* zrows) {
  double* v = malloc((size_t)t->n_inst * sizeof(double));
  for (int i = 0; i < t->n_inst; ++i) v[i] = instance_energy(t, i, zrows);
  double E = ora_reduce_tree(t->n_inst, v);
  free(v);
  return E;
}

/* The product p of one program entry's factors, exactly as site_coeffs! forms
 * it (energy.jl:113-136 / gpu_sweep.jl:55-81). */
static double entry_product(const ora_tiled* t, int j, int e, const double* zrows) {
  int nlm = t->nlm;
  int col = t->site_col.a[j - 1];
  if (col > 0) return ZR(zrows, nlm, t->pent_row.a[e - 1], col);
  if (col < 0)
    return ZR(zrows, nlm, t->pent_row.a[e - 1], -col) *
           ZR(zrows, nlm, t->pent_row2.a[e - 1], t->site_col2.a[j - 1]);
  int off = t->inst_ptr.a[t->site_inst.a[j - 1] - 1] - 1;
  double p = 1.0;
  for (int f = t->sfac_ptr.a[e - 1]; f < t->sfac_ptr.a[e]; ++f)
    p *= ZR(zrows, nlm, t->sfac_row.a[f - 1], t->inst_sites.a[off + t->sfac_slot.a[f - 1] - 1]);
  return p;
}

void ora_site_coeffs(const ora_tiled* t, int s, const double* zrows, double* c) {
  for (int j = t->site_ptr.a[s - 1]; j < t->site_ptr.a[s]; ++j) {
    int pid = t->site_prog.a[j - 1];
    for (int e = t->sprog_ptr.a[pid - 1]; e < t->sprog_ptr.a[pid]; ++e) {
      double p = entry_product(t, j, e, zrows);
      if (p == 0.0) continue;
      c[t->sent_tgt.a[e - 1] - 1] += t->sent_w.a[e - 1] * p;
    }
  }
}

/* The all-site energy gradient G[s] = dE/de_s, tangent-projected (the
 * reference's _gradient_lane_ref!, gpu/gpu_gradient.jl:242-279, with
 * _entry_walk_grad :92-126): the adjacency programs of site s are dealt to ws
 * lanes (lane l takes j = site_ptr[s] + l, + ws, ...); within a program each
 * entry with a nonzero product p (the sweep's skips) adds (w p) grad Z_tgt(e_s)
 * to the lane's 3-vector, w p formed first, then one multiply and one add per
 * component; the lane partials are added in lane order.  config: 3 x n_sites;
 * zrows the rows of config; G: 3 x n_sites. */
void ora_energy_gradient(const ora_tiled* t, const double* config, const double* zrows, int ws, double* G) {
  double grow[3 * 49];
  double* part = malloc((size_t)ws * 3 * sizeof(double));
  for (int s = 1; s <= t->n_sites; ++s) {
    ora_grad_zlm_row(config + 3 * (size_t)(s - 1), t->lmax, grow);
    for (int lane = 0; lane < ws; ++lane) {
      double a[3] = {0.0, 0.0, 0.0};
      for (int j = t->site_ptr.a[s - 1] + lane; j < t->site_ptr.a[s]; j += ws) {
        int pid = t->site_prog.a[j - 1];
        for (int e = t->sprog_ptr.a[pid - 1]; e < t->sprog_ptr.a[pid]; ++e) {
          double p = entry_product(t, j, e, zrows);
          if (p == 0.0) continue;
          int k = t->sent_tgt.a[e - 1] - 1;
          double w = t->sent_w.a[e - 1] * p;
          a[0] = a[0] + w * grow[3 * k];
          a[1] = a[1] + w * grow[3 * k + 1];
          a[2] = a[2] + w * grow[3 * k + 2];
        }
      }
      part[3 * lane] = a[0];
      part[3 * lane + 1] = a[1];
      part[3 * lane + 2] = a[2];
    }
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int lane = 0; lane < ws; ++lane) {
      gx += part[3 * lane];
      gy += part[3 * lane + 1];
      gz += part[3 * lane + 2];
    }
    G[3 * (size_t)(s - 1)] = gx;
    G[3 * (size_t)(s - 1) + 1] = gy;
    G[3 * (size_t)(s - 1) + 2] = gz;
  }
  free(part);
}

double ora_delta_energy(int nlm, const double* c, const double* zold, const double* znew) {
  double dE = 0.0;
  for (int k = 0; k < nlm; ++k) {
    double ck = c[k];
    if (ck == 0.0) continue;
    dE += ck * (znew[k] - zold[k]);
  }
  return dE;
}

/* ======================================================================
 * Keyed proposal — src/gpu/gpu_sweep.jl:32-41, src/updates.jl:20,25-29.
 * normalize(v) = v / sqrt((x*x + y*y) + z*z) component-wise.
 * ====================================================================== */
static void rotate3(const double e[3], const double ax[3], double th, double out[3]) {
  double s, c;
  ora_sincos(th, &s, &c);
  double cr[3] = {ax[1] * e[2] - ax[2] * e[1], ax[2] * e[0] - ax[0] * e[2],
                  ax[0] * e[1] - ax[1] * e[0]};
  double d = (ax[0] * e[0] + ax[1] * e[1]) + ax[2] * e[2];
  double g = (1.0 - c) * d;
  for (int i = 0; i < 3; ++i) out[i] = (c * e[i] + s * cr[i]) + g * ax[i];
}

void ora_keyed_proposal(uint64_t seed, int32_t site, int32_t sweep, const double e[3],
                        double step, double e2[3], double* u_acc) {
  uint32_t b[4];
  ora_philox_block(seed, site, sweep, 0u, 0u, b);
  double u_flip = ora_uniform(b[0], b[1]);
  *u_acc = ora_uniform(b[2], b[3]);
  if (u_flip < 0.2) { e2[0] = -e[0]; e2[1] = -e[1]; e2[2] = -e[2]; return; }
  double n12[2], n34[2];
  ora_philox_block(seed, site, sweep, 1u, 0u, b);
  ora_normal2(b, n12);
  ora_philox_block(seed, site, sweep, 2u, 0u, b);
  ora_normal2(b, n34);
  double nn = sqrt((n12[0] * n12[0] + n12[1] * n12[1]) + n34[0] * n34[0]);
  double ax[3] = {n12[0] / nn, n12[1] / nn, n34[0] / nn};
  rotate3(e, ax, step * n34[1], e2);
}

/* ======================================================================
 * Device walk contract.  The reference kernel strides ws lanes over the CSR
 * adjacency entries and folds the lane partials serially
 * (gpu_sweep.jl:48-85,132-135).  The device kernel here instead walks, per site,
 * the adjacency entries in ascending program id (a per-sublattice invariant
 * order of a tiled Hamiltonian), flattens (entry, program entry) pairs and deals
 * them round-robin over `lanes` lanes; each lane accumulates
 * a = fma(w*p, dz[tgt], a) for p != 0 (dz = znew - zold), and the lane partials
 * are folded by a shuffle-down tree (offsets lanes/2 ... 1).
 * ====================================================================== */
static int* sorted_adjacency(const ora_tiled* t, int s, int* n) {
  int lo = t->site_ptr.a[s - 1], hi = t->site_ptr.a[s];
  *n = hi - lo;
  int* js = malloc((size_t)(*n + 1) * sizeof(int));
  for (int i = 0; i < *n; ++i) js[i] = lo + i;
  for (int i = 1; i < *n; ++i) { /* insertion sort by pid (pids unique per site) */
    int x = js[i], k = i - 1;
    while (k >= 0 && t->site_prog.a[js[k] - 1] > t->site_prog.a[x - 1]) { js[k + 1] = js[k]; --k; }
    js[k + 1] = x;
  }
  return js;
}

/* lanes: a power of two up to ORA_MAX_LANES */
#define ORA_MAX_LANES 256

static void lane_fold(double* a, int lanes) {
  for (int off = lanes / 2; off >= 1; off >>= 1)
    for (int l = 0; l < off; ++l) a[l] = a[l] + a[l + off];
}

double ora_walk_delta(const ora_tiled* t, int s, const double* zrows, const double* znew,
                      int lanes) {
  int nlm = t->nlm;
  double dz[64];
  const double* zold = zrows + (size_t)(s - 1) * nlm;
  for (int k = 0; k < nlm; ++k) dz[k] = znew[k] - zold[k];
  double a[ORA_MAX_LANES] = {0};
  int n, i = 0;
  int* js = sorted_adjacency(t, s, &n);
  for (int q = 0; q < n; ++q) {
    int j = js[q], pid = t->site_prog.a[j - 1];
    for (int e = t->sprog_ptr.a[pid - 1]; e < t->sprog_ptr.a[pid]; ++e, ++i) {
      double p = entry_product(t, j, e, zrows);
      if (p == 0.0) continue;
      int l = i % lanes;
      a[l] = fma(t->sent_w.a[e - 1] * p, dz[t->sent_tgt.a[e - 1] - 1], a[l]);
    }
  }
  free(js);
  lane_fold(a, lanes);
  return a[0];
}

void ora_walk_coeffs(const ora_tiled* t, int s, const double* zrows, int lanes, double* c) {
  int nlm = t->nlm;
  double a[ORA_MAX_LANES][49];
  memset(a, 0, sizeof a);
  int n, i = 0;
  int* js = sorted_adjacency(t, s, &n);
  for (int q = 0; q < n; ++q) {
    int j = js[q], pid = t->site_prog.a[j - 1];
    for (int e = t->sprog_ptr.a[pid - 1]; e < t->sprog_ptr.a[pid]; ++e, ++i) {
      double p = entry_product(t, j, e, zrows);
      if (p == 0.0) continue;
      int l = i % lanes, k = t->sent_tgt.a[e - 1] - 1;
      a[l][k] = a[l][k] + t->sent_w.a[e - 1] * p;
    }
  }
  free(js);
  for (int k = 0; k < nlm; ++k) {
    double v[ORA_MAX_LANES];
    for (int l = 0; l < lanes; ++l) v[l] = a[l][k];
    lane_fold(v, lanes);
    c[k] = v[0];
  }
}

/* One device-contract Metropolis sweep (gpu_sweep.jl:158-190 semantics):
 * proposals keyed by (seed, site, sweep); colour-serial accept; per-site dE/acc
 * staged in colour order; the sweep's ΔE reduced by the fixed tree. */
int64_t ora_keyed_sweep(const ora_tiled* t, double* config, double* zrows, double* energy,
                        double beta, double step, uint64_t seed, int32_t sweep, int lanes,
                        double* dE_q, int32_t* acc_q) {
  int nlm = t->nlm;
  double znew[49], e2[3], u;
  int64_t nacc = 0;
  for (int q = 0; q < t->n_active; ++q) {
    int s = t->color_sites.a[q];
    double* e = config + 3 * (size_t)(s - 1);
    ora_keyed_proposal(seed, s, sweep, e, step, e2, &u);
    ora_zlm_row(e2, t->lmax, znew);
    double dE = ora_walk_delta(t, s, zrows, znew, lanes);
    if (dE <= 0.0 || u < ora_exp(-beta * dE)) {
      e[0] = e2[0]; e[1] = e2[1]; e[2] = e2[2];
      memcpy(zrows + (size_t)(s - 1) * nlm, znew, (size_t)nlm * sizeof(double));
      dE_q[q] = dE; acc_q[q] = 1; ++nacc;
    } else {
      dE_q[q] = 0.0; acc_q[q] = 0;
    }
  }
  *energy = *energy + ora_reduce_tree(t->n_active, dE_q);
  return nacc;
}

/* Device-contract overrelaxation sweep (updates.jl:66-87 move, colour order of
 * :231-290).  The accept uniform is keyed: block (seed, site, sweep, 16+pass, 0),
 * words 1-2 — the reference's CPU OR draws from the site stream instead. */
int64_t ora_keyed_or_sweep(const ora_tiled* t, double* config, double* zrows, double* energy,
                           double beta, uint64_t seed, int32_t sweep, int32_t or_pass,
                           int lanes, double* dE_q, int32_t* acc_q, int64_t* natt) {
  int nlm = t->nlm;
  double c[49], znew[49];
  int64_t nacc = 0;
  *natt = 0;
  for (int q = 0; q < t->n_active; ++q) {
    int s = t->color_sites.a[q];
    dE_q[q] = 0.0; acc_q[q] = 0;
    if (!t->site_has_l1.a[s - 1]) continue;
    ora_walk_coeffs(t, s, zrows, lanes, c);
    double h[3] = {c[3], c[1], c[2]};
    double hn = sqrt((h[0] * h[0] + h[1] * h[1]) + h[2] * h[2]);
    if (hn < 1e-12) continue;
    ++*natt;
    double ax[3] = {h[0] / hn, h[1] / hn, h[2] / hn};
    double* e = config + 3 * (size_t)(s - 1);
    double d = (e[0] * ax[0] + e[1] * ax[1]) + e[2] * ax[2];
    double tt = 2.0 * d;
    double e2[3] = {tt * ax[0] - e[0], tt * ax[1] - e[1], tt * ax[2] - e[2]};
    ora_zlm_row(e2, t->lmax, znew);
    double* zold = zrows + (size_t)(s - 1) * nlm;
    double dE = ora_delta_energy(nlm, c, zold, znew);
    uint32_t b[4];
    ora_philox_block(seed, s, sweep, 16u + (uint32_t)or_pass, 0u, b);
    double u = ora_uniform(b[0], b[1]);
    if (dE <= 0.0 || u < ora_exp(-beta * dE)) {
      e[0] = e2[0]; e[1] = e2[1]; e[2] = e2[2];
      memcpy(zold, znew, (size_t)nlm * sizeof(double));
      dE_q[q] = dE; acc_q[q] = 1; ++nacc;
    }
  }
  *energy = *energy + ora_reduce_tree(t->n_active, dE_q);
  return nacc;
}

/* Renormalisation with full re-anchoring — src/state.jl:168-185, with the
 * total energy reduced by the device tree.  Returns the drift. */
double ora_renormalize(const ora_tiled* t, double* config, double* zrows, double* energy) {
  for (int s = 0; s < t->n_sites; ++s) {
    if (!t->site_active.a[s]) continue;
    double* e = config + 3 * (size_t)s;
    double n = sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]);
    e[0] = e[0] / n; e[1] = e[1] / n; e[2] = e[2] / n;
    ora_zlm_row(e, t->lmax, zrows + (size_t)s * t->nlm);
  }
  double E = ora_total_energy_tree(t, zrows);
  double drift = fabs(*energy - E);
  *energy = E;
  return drift;
}

/* ======================================================================
 * Reference CPU path — src/updates.jl:37-58 (_attempt_metro!), :66-87
 * (_attempt_or!), :92-98 (_reduce_dE), :157-206 / :231-290 (colour sweeps,
 * strided tasks + barrier); src/state.jl:35-46 (per-site streams), :80-86
 * (_random_unit).  Per-site streams are xoshiro256++ (Blackman & Vigna),
 * seeded by splitmix64, and normals come from a cached Box-Muller pair:
 * statistically equivalent to Julia's Xoshiro/randn, not bit-equal.
 * ====================================================================== */
typedef struct { uint64_t s[4]; int have; double spare; } xrng;

static uint64_t splitmix64(uint64_t* x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t xnext(xrng* r) {
  uint64_t* s = r->s;
  uint64_t res = rotl(s[0] + s[3], 23) + s[0];
  uint64_t t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
  s[2] ^= t; s[3] = rotl(s[3], 45);
  return res;
}
static void xseed(xrng* r, uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
  r->have = 0;
}
static double xrand(xrng* r) { return (double)(xnext(r) >> 11) * 0x1p-53; }
static double xrandn(xrng* r) {
  if (r->have) { r->have = 0; return r->spare; }
  double u1 = ((double)(xnext(r) >> 11) + 0.5) * 0x1p-53;
  double u2 = (double)(xnext(r) >> 11) * 0x1p-53;
  double rr = sqrt(-2.0 * ora_log(u1));
  double s, c;
  ora_sincos(6.283185307179586 * u2, &s, &c);
  r->spare = rr * s;
  r->have = 1;
  return rr * c;
}

struct ora_cpu_chain {
  const ora_tiled* t;
  double *config, *zrows, *dE;
  double energy, step, max_drift;
  int64_t acc_metro, att_metro, acc_or, att_or;
  xrng* site_rngs;
};

ora_cpu_chain* ora_cpu_chain_create(const ora_tiled* t, const double* config, uint64_t seed,
                                    double step) {
  ora_cpu_chain* c = calloc(1, sizeof *c);
  c->t = t;
  size_t N = (size_t)t->n_sites;
  c->config = malloc(3 * N * sizeof(double));
  memcpy(c->config, config, 3 * N * sizeof(double));
  c->zrows = malloc(N * (size_t)t->nlm * sizeof(double));
  ora_zrows(t, c->config, c->zrows);
  c->energy = ora_total_energy(t, c->zrows);
  c->dE = calloc(N, sizeof(double));
  c->step = step;
  c->site_rngs = malloc(N * sizeof(xrng));
  xrng master;
  xseed(&master, seed);
  for (size_t s = 0; s < N; ++s) xseed(&c->site_rngs[s], xnext(&master));
  return c;
}

void ora_cpu_chain_free(ora_cpu_chain* c) {
  if (!c) return;
  free(c->config); free(c->zrows); free(c->dE); free(c->site_rngs); free(c);
}

static int attempt_metro(ora_cpu_chain* ch, double beta, int s, double* cbuf, double* znew) {
  const ora_tiled* t = ch->t;
  int nlm = t->nlm;
  xrng* rng = &ch->site_rngs[s - 1];
  memset(cbuf, 0, (size_t)nlm * sizeof(double));
  ora_site_coeffs(t, s, ch->zrows, cbuf);
  double* e = ch->config + 3 * (size_t)(s - 1);
  double e2[3];
  if (xrand(rng) < 0.2) {
    e2[0] = -e[0]; e2[1] = -e[1]; e2[2] = -e[2];
  } else {
    double v[3], n;
    do { /* _random_unit, state.jl:80-86 */
      v[0] = xrandn(rng); v[1] = xrandn(rng); v[2] = xrandn(rng);
      n = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    } while (!(n > 1e-12));
    double ax[3] = {v[0] / n, v[1] / n, v[2] / n};
    rotate3(e, ax, ch->step * xrandn(rng), e2);
  }
  ora_zlm_row(e2, t->lmax, znew);
  double* zold = ch->zrows + (size_t)(s - 1) * nlm;
  double dE = ora_delta_energy(nlm, cbuf, zold, znew);
  if (dE <= 0.0 || xrand(rng) < ora_exp(-beta * dE)) {
    e[0] = e2[0]; e[1] = e2[1]; e[2] = e2[2];
    memcpy(zold, znew, (size_t)nlm * sizeof(double));
    ch->dE[s - 1] = dE;
    return 1;
  }
  return 0;
}

static int attempt_or(ora_cpu_chain* ch, double beta, int s, double* cbuf, double* znew,
                      int* att) {
  const ora_tiled* t = ch->t;
  int nlm = t->nlm;
  xrng* rng = &ch->site_rngs[s - 1];
  memset(cbuf, 0, (size_t)nlm * sizeof(double));
  ora_site_coeffs(t, s, ch->zrows, cbuf);
  double h[3] = {cbuf[3], cbuf[1], cbuf[2]};
  double hn = sqrt((h[0] * h[0] + h[1] * h[1]) + h[2] * h[2]);
  *att = 0;
  if (hn < 1e-12) return 0;
  *att = 1;
  double ax[3] = {h[0] / hn, h[1] / hn, h[2] / hn};
  double* e = ch->config + 3 * (size_t)(s - 1);
  double d = (e[0] * ax[0] + e[1] * ax[1]) + e[2] * ax[2];
  double tt = 2.0 * d;
  double e2[3] = {tt * ax[0] - e[0], tt * ax[1] - e[1], tt * ax[2] - e[2]};
  ora_zlm_row(e2, t->lmax, znew);
  double* zold = ch->zrows + (size_t)(s - 1) * nlm;
  double dE = ora_delta_energy(nlm, cbuf, zold, znew);
  if (dE <= 0.0 || xrand(rng) < ora_exp(-beta * dE)) {
    e[0] = e2[0]; e[1] = e2[1]; e[2] = e2[2];
    memcpy(zold, znew, (size_t)nlm * sizeof(double));
    ch->dE[s - 1] = dE;
    return 1;
  }
  return 0;
}

typedef struct {
  ora_cpu_chain* ch;
  double beta;
  int tid, nt, kind;
  pthread_barrier_t* bar;
  int64_t acc, att;
} cpu_task;

static void* cpu_worker(void* arg) {
  cpu_task* w = arg;
  const ora_tiled* t = w->ch->t;
  double cbuf[49], znew[49];
  for (int c = 0; c < t->n_colors; ++c) {
    int q1 = t->color_ptr.a[c + 1] - 1;
    for (int q = t->color_ptr.a[c] + w->tid; q <= q1; q += w->nt) {
      int s = t->color_sites.a[q - 1];
      if (w->kind == 0) {
        w->acc += attempt_metro(w->ch, w->beta, s, cbuf, znew);
      } else {
        if (!t->site_has_l1.a[s - 1]) continue;
        int at = 0;
        w->acc += attempt_or(w->ch, w->beta, s, cbuf, znew, &at);
        w->att += at;
      }
    }
    if (w->nt > 1) pthread_barrier_wait(w->bar);
  }
  return NULL;
}

static int64_t cpu_sweep(ora_cpu_chain* ch, double beta, int nthreads, int kind, int64_t* natt) {
  const ora_tiled* t = ch->t;
  memset(ch->dE, 0, (size_t)t->n_sites * sizeof(double));
  if (nthreads < 1) nthreads = 1;
  cpu_task tasks[256];
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  pthread_barrier_t bar;
  if (nthreads > 1) pthread_barrier_init(&bar, NULL, (unsigned)nthreads);
  for (int i = 0; i < nthreads; ++i) {
    tasks[i].ch = ch; tasks[i].beta = beta; tasks[i].tid = i; tasks[i].nt = nthreads;
    tasks[i].kind = kind; tasks[i].bar = &bar; tasks[i].acc = 0; tasks[i].att = 0;
  }
  if (nthreads == 1) {
    cpu_worker(&tasks[0]);
  } else {
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, cpu_worker, &tasks[i]);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    pthread_barrier_destroy(&bar);
  }
  int64_t acc = 0, att = 0;
  for (int i = 0; i < nthreads; ++i) { acc += tasks[i].acc; att += tasks[i].att; }
  double dsum = 0.0; /* _reduce_dE: serial in colour order */
  for (int q = 0; q < t->n_active; ++q) dsum += ch->dE[t->color_sites.a[q] - 1];
  ch->energy += dsum;
  *natt = att;
  return acc;
}

int64_t ora_cpu_metropolis_sweep(ora_cpu_chain* ch, double beta, int nthreads) {
  int64_t natt;
  int64_t acc = cpu_sweep(ch, beta, nthreads, 0, &natt);
  ch->acc_metro += acc;
  ch->att_metro += ch->t->n_active;
  return acc;
}

int64_t ora_cpu_or_sweep(ora_cpu_chain* ch, double beta, int nthreads) {
  int64_t natt;
  int64_t acc = cpu_sweep(ch, beta, nthreads, 1, &natt);
  ch->acc_or += acc;
  ch->att_or += natt;
  return acc;
}

/* state.jl:168-185 with the reference's serial total energy. */
double ora_cpu_renormalize(ora_cpu_chain* ch) {
  const ora_tiled* t = ch->t;
  for (int s = 0; s < t->n_sites; ++s) {
    if (!t->site_active.a[s]) continue;
    double* e = ch->config + 3 * (size_t)s;
    double n = sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]);
    e[0] = e[0] / n; e[1] = e[1] / n; e[2] = e[2] / n;
    ora_zlm_row(e, t->lmax, ch->zrows + (size_t)s * t->nlm);
  }
  double E = ora_total_energy(t, ch->zrows);
  double drift = fabs(ch->energy - E);
  if (drift > ch->max_drift) ch->max_drift = drift;
  ch->energy = E;
  return drift;
}

void ora_cpu_state(const ora_cpu_chain* ch, double* config, double* zrows, double* energy,
                   int64_t counters[4]) {
  size_t N = (size_t)ch->t->n_sites;
  if (config) memcpy(config, ch->config, 3 * N * sizeof(double));
  if (zrows) memcpy(zrows, ch->zrows, N * (size_t)ch->t->nlm * sizeof(double));
  if (energy) *energy = ch->energy;
  if (counters) ora_cpu_counters(ch, counters);
}
double ora_cpu_energy(const ora_cpu_chain* ch) { return ch->energy; }
void ora_cpu_mean_spin(const ora_cpu_chain* ch, double m[3]) {
  const ora_tiled* t = ch->t;
  double a[3] = {0, 0, 0};
  for (int s = 0; s < t->n_sites; ++s) {
    if (!t->site_active.a[s]) continue;
    for (int i = 0; i < 3; ++i) a[i] += ch->config[3 * (size_t)s + i];
  }
  for (int i = 0; i < 3; ++i) m[i] = a[i] / (double)t->n_active;
}
double ora_cpu_step(const ora_cpu_chain* ch) { return ch->step; }
void ora_cpu_set_step(ora_cpu_chain* ch, double step) { ch->step = step; }
void ora_cpu_reset_counters(ora_cpu_chain* ch) {
  ch->acc_metro = ch->att_metro = ch->acc_or = ch->att_or = 0;
}
void ora_cpu_counters(const ora_cpu_chain* ch, int64_t c[4]) {
  c[0] = ch->acc_metro; c[1] = ch->att_metro; c[2] = ch->acc_or; c[3] = ch->att_or;
}

/* ======================================================================
 * Learned delayed acceptance (an extension of the library; no reference code).
 * Descriptor of site s: its neighbour sites in order of first appearance in
 * the pid-ordered walk (factor order as in site_coeffs!), features
 * [e.e_n ..., (e.e_n)^2 ..., one per triplet: e.(e_x x e_y), one per body-4
 * cluster: (e.e_x)(e_y.e_z) + (e.e_y)(e_x.e_z) + (e.e_z)(e_x.e_y), and the
 * three spin components when a body-1 term touches s].  Features are formed in
 * FP64 and rounded to FP32; the MLP (F -> 64 -> 64 -> 1, tanh, + linear path)
 * runs in FP32 in the device's operation order.  Two-stage acceptance
 * (Christen & Fox 2005): stage 1 on dE_s, stage 2 on dE - dE_s with the slot-3
 * keyed uniform.
 * ====================================================================== */
#define SUR_H 64

static double dot3d(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

static int site_factor_sites(const ora_tiled* t, int j, int* fs) {
  int col = t->site_col.a[j - 1];
  if (col > 0) { fs[0] = col; return 1; }
  if (col < 0) { fs[0] = -col; fs[1] = t->site_col2.a[j - 1]; return 2; }
  int inst = t->site_inst.a[j - 1];
  int off = t->inst_ptr.a[inst - 1] - 1;
  int body = t->inst_ptr.a[inst] - t->inst_ptr.a[inst - 1];
  int own = t->site_slot.a[j - 1], n = 0;
  for (int k = 1; k <= body; ++k) if (k != own) fs[n++] = t->inst_sites.a[off + k - 1];
  return n;
}

/* Features of site s for spin e (3) against config; returns F.  Needs
 * lmax >= 1 (the neighbour spins are read back from their l = 1 rows). */
int ora_sur_features(const ora_tiled* t, const double* config, int s, const double* e, float* out) {
  int n, nb[256], nnb = 0, spec[512][4], nspec = 0, body1 = 0;
  int* js = sorted_adjacency(t, s, &n);
  for (int q = 0; q < n; ++q) {
    int fs[8], idx[8];
    int nf = site_factor_sites(t, js[q], fs);
    for (int f = 0; f < nf; ++f) {
      int k = 0;
      while (k < nnb && nb[k] != fs[f]) ++k;
      if (k == nnb) nb[nnb++] = fs[f];
      idx[f] = k;
    }
    int col = t->site_col.a[js[q] - 1];
    if (col < 0) { spec[nspec][0] = 1; spec[nspec][1] = idx[0]; spec[nspec][2] = idx[1]; ++nspec; }
    else if (col == 0 && nf == 3) { spec[nspec][0] = 2; spec[nspec][1] = idx[0]; spec[nspec][2] = idx[1]; spec[nspec][3] = idx[2]; ++nspec; }
    else if (col == 0 && nf == 0) body1 = 1;
  }
  free(js);
  if (body1) for (int x = 0; x < 3; ++x) { spec[nspec][0] = 3; spec[nspec][1] = x; ++nspec; }
  /* neighbour spins read back from their tesseral rows: Z_{1,-1} = N1 y,
   * Z_{1,0} = N1 z, Z_{1,1} = N1 x, so e_n = (Z[3], Z[1], Z[2]) * (1 / N1) */
  static const double inv_n1 = 0x1.05f8bd37c0e62p+1;
  double Es[256][3], zr[49];
  const double* E[256];
  for (int k = 0; k < nnb; ++k) {
    ora_zlm_row(config + 3 * (size_t)(nb[k] - 1), t->lmax, zr);
    Es[k][0] = zr[3] * inv_n1;
    Es[k][1] = zr[1] * inv_n1;
    Es[k][2] = zr[2] * inv_n1;
    E[k] = Es[k];
  }
  int F = 0;
  for (int k = 0; k < nnb; ++k) out[F++] = (float)dot3d(e, E[k]);
  for (int k = 0; k < nnb; ++k) { double a = dot3d(e, E[k]); out[F++] = (float)(a * a); }
  for (int q = 0; q < nspec; ++q) {
    if (spec[q][0] == 1) {
      const double *x = E[spec[q][1]], *y = E[spec[q][2]];
      double cr[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
      out[F++] = (float)dot3d(e, cr);
    } else if (spec[q][0] == 2) {
      const double *x = E[spec[q][1]], *y = E[spec[q][2]], *z = E[spec[q][3]];
      out[F++] = (float)((dot3d(e, x) * dot3d(y, z) + dot3d(e, y) * dot3d(x, z)) + dot3d(e, z) * dot3d(x, y));
    } else {
      out[F++] = (float)e[spec[q][1]];
    }
  }
  return F;
}

/* The surrogate's activation in IEEE single ops (the device's det_tanhf): odd
 * Taylor polynomial of tanh through x^11 for |x| < 1/4; (1 - t)/(1 + t) with
 * t = e^{-2|x|} for |x| <= 10, the exponential by Cody-Waite reduction
 * (ln 2 = 0x1.63p-1 + -0x1.bd0106p-13), degree-7 Taylor polynomial and an exact
 * 2^k; 1 beyond.  Pinned against tanh (<= 2.2 ulp), oddness, monotonicity and
 * the two zeros by the oracle's tests. */
static const float TANH_C[5] = {-0x1.555556p-2f, 0x1.111112p-3f, -0x1.ba1ba2p-5f, 0x1.664f48p-6f,
                                -0x1.226e36p-7f}; /* x^3 .. x^11: -1/3, 2/15, -17/315, 62/2835, -1382/155925 */
static const float EXPF_C[8] = {1.0f, 1.0f, 0.5f, 0x1.555556p-3f, 0x1.555556p-5f, 0x1.111112p-7f,
                                0x1.6c16c2p-10f, 0x1.a01a02p-13f}; /* 1/n!, n = 0..7 */

static float ora_tanhf(float z) {
  float ax = fabsf(z), r;
  if (ax < 0.25f) {
    float x2 = ax * ax, p = TANH_C[4];
    for (int k = 3; k >= 0; --k) p = p * x2 + TANH_C[k];
    r = ax + ax * (x2 * p);
  } else if (ax > 10.0f) {
    r = 1.0f;
  } else {
    float y = -2.0f * ax;
    float k = rintf(y * 0x1.715476p+0f);
    float q = (y - k * 0x1.63p-1f) - k * -0x1.bd0106p-13f;
    float p = EXPF_C[7];
    for (int n = 6; n >= 0; --n) p = p * q + EXPF_C[n];
    uint32_t b = (uint32_t)((int)k + 127) << 23;
    float s;
    memcpy(&s, &b, 4);
    float t = p * s;
    r = (1.0f - t) / (1.0f + t);
  }
  return z < 0.0f ? -r : r;
}

float ora_test_tanhf(float z) { return ora_tanhf(z); }

/* params: [W1 (64 x F), b1 (64), W2 (64 x 64), b2 (64), w3 (64), v (F)].
 * Hidden pre-activations are fused multiply-add chains in ascending input
 * index starting from the bias, over the input zero-padded to a multiple of 4
 * (F4): z = fma(x[F4-1], w[F4-1], ... fma(x[0], w[0], b)).  This is what the
 * device's matrix-core tiles (v_mfma_f32_16x16x4f32: four fused multiply-adds
 * in k order per instruction, instructions in ascending k) compute.  The head
 * w3 . h2 + v . phi keeps separate multiplies and adds, folded over 64 lanes
 * (lane j: w3[j] h2[j] + sum over f = j, j+64, ... of v[f] phi[f]). */
double ora_sur_delta(int F, const float* params, double scale, const float* phiA, const float* phiB) {
  const float* W1 = params;
  const float* b1 = W1 + SUR_H * F;
  const float* W2 = b1 + SUR_H;
  const float* b2 = W2 + SUR_H * SUR_H;
  const float* w3 = b2 + SUR_H;
  const float* v = w3 + SUR_H;
  const int F4 = (F + 3) & ~3;
  float xA[520], xB[520];
  for (int f = 0; f < F4; ++f) {
    xA[f] = f < F ? phiA[f] : 0.0f;
    xB[f] = f < F ? phiB[f] : 0.0f;
  }
  float h1A[SUR_H], h1B[SUR_H], tA[SUR_H], tB[SUR_H];
  for (int j = 0; j < SUR_H; ++j) {
    float zA = b1[j], zB = b1[j];
    for (int f = 0; f < F4; ++f) {
      const float w = f < F ? W1[j * F + f] : 0.0f;
      zA = fmaf(xA[f], w, zA);
      zB = fmaf(xB[f], w, zB);
    }
    h1A[j] = ora_tanhf(zA);
    h1B[j] = ora_tanhf(zB);
  }
  for (int j = 0; j < SUR_H; ++j) {
    float zA = b2[j], zB = b2[j];
    for (int k = 0; k < SUR_H; ++k) {
      const float w = W2[j * SUR_H + k];
      zA = fmaf(h1A[k], w, zA);
      zB = fmaf(h1B[k], w, zB);
    }
    float oA = w3[j] * ora_tanhf(zA), oB = w3[j] * ora_tanhf(zB);
    float sA = 0.0f, sB = 0.0f;
    for (int f = j; f < F; f += SUR_H) {
      sA = sA + v[f] * phiA[f];
      sB = sB + v[f] * phiB[f];
    }
    tA[j] = oA + sA;
    tB[j] = oB + sB;
  }
  for (int off = SUR_H / 2; off >= 1; off >>= 1)
    for (int l = 0; l < off; ++l) { tA[l] = tA[l] + tA[l + off]; tB[l] = tB[l] + tB[l + off]; }
  return ((double)tB[0] - (double)tA[0]) * scale;
}

/* The int8 surrogate (qp: the quantised network blob, sce_sur_get_qparams:
 * V1 (64 x F), r1, b1, V2 (64 x 64), r2, b2, w3 (64 each), v (F), input
 * scales (F)).  Inputs u_f = clamp(rint(phi_f * s_f), -127, 127); a layer's
 * pre-activation fmaf((float)acc, r_j, b_j) of the exact integer dot product
 * acc; layer 1's activations enter layer 2 as rint(127 tanh(z)); the head is
 * ora_sur_delta's. */
double ora_sur_delta_q(int F, const float* qp, double scale, const float* phiA, const float* phiB) {
  const float* V1 = qp;
  const float* r1 = V1 + SUR_H * F;
  const float* b1 = r1 + SUR_H;
  const float* V2 = b1 + SUR_H;
  const float* r2 = V2 + SUR_H * SUR_H;
  const float* b2 = r2 + SUR_H;
  const float* w3 = b2 + SUR_H;
  const float* v = w3 + SUR_H;
  const float* qin = v + F;
  int uA[512], uB[512], hA[SUR_H], hB[SUR_H];
  for (int f = 0; f < F; ++f) {
    float xa = rintf(phiA[f] * qin[f]), xb = rintf(phiB[f] * qin[f]);
    uA[f] = (int)fminf(127.0f, fmaxf(-127.0f, xa));
    uB[f] = (int)fminf(127.0f, fmaxf(-127.0f, xb));
  }
  for (int j = 0; j < SUR_H; ++j) {
    int64_t aA = 0, aB = 0;
    for (int f = 0; f < F; ++f) {
      aA += (int64_t)V1[j * F + f] * uA[f];
      aB += (int64_t)V1[j * F + f] * uB[f];
    }
    hA[j] = (int)rintf(127.0f * ora_tanhf(fmaf((float)aA, r1[j], b1[j])));
    hB[j] = (int)rintf(127.0f * ora_tanhf(fmaf((float)aB, r1[j], b1[j])));
  }
  float tA[SUR_H], tB[SUR_H];
  for (int j = 0; j < SUR_H; ++j) {
    int64_t aA = 0, aB = 0;
    for (int k = 0; k < SUR_H; ++k) {
      aA += (int64_t)V2[j * SUR_H + k] * hA[k];
      aB += (int64_t)V2[j * SUR_H + k] * hB[k];
    }
    float oA = w3[j] * ora_tanhf(fmaf((float)aA, r2[j], b2[j]));
    float oB = w3[j] * ora_tanhf(fmaf((float)aB, r2[j], b2[j]));
    float sA = 0.0f, sB = 0.0f;
    for (int f = j; f < F; f += SUR_H) {
      sA = sA + v[f] * phiA[f];
      sB = sB + v[f] * phiB[f];
    }
    tA[j] = oA + sA;
    tB[j] = oB + sB;
  }
  for (int off = SUR_H / 2; off >= 1; off >>= 1)
    for (int l = 0; l < off; ++l) { tA[l] = tA[l] + tA[l + off]; tB[l] = tB[l] + tB[l + off]; }
  return ((double)tB[0] - (double)tA[0]) * scale;
}

/* One keyed delayed-acceptance sweep (device contract, 64 lanes).  params[a],
 * F[a], scale[a] per sublattice a = (site-1) % n_cell_atoms.  Returns nacc;
 * *n_stage1 = proposals that passed the surrogate stage. */
static int64_t da_sweep(const ora_tiled* t, double* config, double* zrows, double* energy, double beta, double step,
                        uint64_t seed, int32_t sweep, const float* const* params, const double* scale, double* dE_q,
                        int32_t* acc_q, int64_t* n_stage1, int quant);

int64_t ora_da_sweep(const ora_tiled* t, double* config, double* zrows, double* energy, double beta,
                     double step, uint64_t seed, int32_t sweep, const float* const* params,
                     const double* scale, double* dE_q, int32_t* acc_q, int64_t* n_stage1) {
  return da_sweep(t, config, zrows, energy, beta, step, seed, sweep, params, scale, dE_q, acc_q, n_stage1, 0);
}

/* The same sweep with stage 1 on the int8 networks (qparams[a]: blobs of
 * ora_sur_delta_q). */
int64_t ora_da_sweep_q(const ora_tiled* t, double* config, double* zrows, double* energy, double beta,
                       double step, uint64_t seed, int32_t sweep, const float* const* qparams,
                       const double* scale, double* dE_q, int32_t* acc_q, int64_t* n_stage1) {
  return da_sweep(t, config, zrows, energy, beta, step, seed, sweep, qparams, scale, dE_q, acc_q, n_stage1, 1);
}

static int64_t da_sweep(const ora_tiled* t, double* config, double* zrows, double* energy, double beta, double step,
                        uint64_t seed, int32_t sweep, const float* const* params, const double* scale, double* dE_q,
                        int32_t* acc_q, int64_t* n_stage1, int quant) {
  int nlm = t->nlm;
  double znew[49], e2[3], u;
  float phiA[512], phiB[512];
  int64_t nacc = 0;
  *n_stage1 = 0;
  /* proposals of the whole sweep first (they depend only on the site's own spin) */
  double* props = malloc((size_t)t->n_active * 4 * sizeof(double));
  for (int q = 0; q < t->n_active; ++q) {
    int s = t->color_sites.a[q];
    ora_keyed_proposal(seed, s, sweep, config + 3 * (size_t)(s - 1), step, props + 4 * q, props + 4 * q + 3);
  }
  for (int q = 0; q < t->n_active; ++q) {
    int s = t->color_sites.a[q];
    double* e = config + 3 * (size_t)(s - 1);
    e2[0] = props[4 * q]; e2[1] = props[4 * q + 1]; e2[2] = props[4 * q + 2]; u = props[4 * q + 3];
    int a = (s - 1) % t->n_cell_atoms;
    int F = ora_sur_features(t, config, s, e, phiA);
    ora_sur_features(t, config, s, e2, phiB);
    double dEs = quant ? ora_sur_delta_q(F, params[a], scale[a], phiA, phiB)
                       : ora_sur_delta(F, params[a], scale[a], phiA, phiB);
    dE_q[q] = 0.0; acc_q[q] = 0;
    if (!(dEs <= 0.0 || u < ora_exp(-beta * dEs))) continue;
    ++*n_stage1;
    ora_zlm_row(e2, t->lmax, znew);
    double dE = ora_walk_delta(t, s, zrows, znew, 64);
    double d2 = dE - dEs;
    uint32_t b[4];
    ora_philox_block(seed, s, sweep, 3u, 0u, b);
    if (d2 <= 0.0 || ora_uniform(b[0], b[1]) < ora_exp(-beta * d2)) {
      e[0] = e2[0]; e[1] = e2[1]; e[2] = e2[2];
      memcpy(zrows + (size_t)(s - 1) * nlm, znew, (size_t)nlm * sizeof(double));
      dE_q[q] = dE; acc_q[q] = 1; ++nacc;
    }
  }
  free(props);
  *energy = *energy + ora_reduce_tree(t->n_active, dE_q);
  return nacc;
}
