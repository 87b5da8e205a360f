// sce_math.h — the arithmetic both the gfx950 kernels and the host library use:
// keyed Philox4x32-10 draws, the project's deterministic exp/log/sincos, the
// tesseral harmonic row, and the keyed Metropolis proposal.
//
// Everything here is __host__ __device__ and uses only IEEE +,-,*,/,sqrt,fma,rint
// and integer ops; the library is compiled with -ffp-contract=off, so host and
// device produce identical bits for identical inputs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SCE_HD __host__ __device__ __forceinline__

namespace sce {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) — src/gpu/philox.jl:13-60.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

SCE_HD u32x4 philox4x32(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xd2511f53u * (uint64_t)c.x;
    uint64_t p1 = (uint64_t)0xcd9e8d57u * (uint64_t)c.z;
    u32x4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9e3779b9u;
    k1 += 0xbb67ae85u;
  }
  return c;
}

// Counter (site, sweep, slot, tag): site/sweep are the reference's 1-based Int32
// coordinates reinterpreted as UInt32; key = (lo32(seed), hi32(seed)).
SCE_HD u32x4 philox_block(uint64_t seed, int32_t site, int32_t sweep, uint32_t slot,
                          uint32_t tag) {
  u32x4 c{(uint32_t)site, (uint32_t)sweep, slot, tag};
  return philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Strictly open (0,1): (top 52 bits + 0.5) * 2^-52 — philox.jl:78-81.
SCE_HD double uniform01(uint32_t hi, uint32_t lo) {
  uint64_t w = ((uint64_t)hi << 32) | (uint64_t)lo;
  return ((double)(w >> 12) + 0.5) * 0x1p-52;
}

// ---------------------------------------------------------------------------
// Deterministic exp / log / sincos: Cody–Waite argument reduction followed by
// truncated Taylor (exp, sin, cos) and atanh (log) series, evaluated by Horner
// in a fixed order.  Accuracy: exp ≤ 1 ulp, log ≤ 2 ulp, sin/cos ≤ 0.5 ulp(1)
// absolute on the ranges the sweep uses.
// ---------------------------------------------------------------------------
SCE_HD double bits_to_d(uint64_t b) { return __builtin_bit_cast(double, b); }
SCE_HD uint64_t d_to_bits(double d) { return __builtin_bit_cast(uint64_t, d); }
SCE_HD double pow2i(int k) { return bits_to_d((uint64_t)(k + 1023) << 52); }

SCE_HD double det_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_inf();
  if (x < -745.2) return 0.0;
  const double n = __builtin_rint(x * 0x1.71547652b82fep+0);
  const double r = (x - n * 0x1.62e42fee00000p-1) - n * 0x1.a39ef35793c76p-33;
  double p = 0x1.6124613a86d09p-33;
  p = p * r + 0x1.1eed8eff8d898p-29;
  p = p * r + 0x1.ae64567f544e4p-26;
  p = p * r + 0x1.27e4fb7789f5cp-22;
  p = p * r + 0x1.71de3a556c734p-19;
  p = p * r + 0x1.a01a01a01a01ap-16;
  p = p * r + 0x1.a01a01a01a01ap-13;
  p = p * r + 0x1.6c16c16c16c17p-10;
  p = p * r + 0x1.1111111111111p-7;
  p = p * r + 0x1.5555555555555p-5;
  p = p * r + 0x1.5555555555555p-3;
  p = p * r + 0x1.0000000000000p-1;
  p = p * r + 1.0;
  p = p * r + 1.0;
  const int k = (int)n;
  if (k > 1023) return p * pow2i(1023) * pow2i(k - 1023);
  if (k < -1022) return p * pow2i(k + 1000) * pow2i(-1000);
  return p * pow2i(k);
}

// tanh in FP32 (the surrogate's activation): an odd Taylor polynomial through
// x^11 below |x| = 1/4; (1 - t)/(1 + t) with t = e^{-2|x|} up to |x| = 10
// (t from a Cody-Waite reduction by ln 2 with a 9-bit high part, the degree-7
// Taylor exponential and an exact 2^k scale; 1 - t >= 0.39, no cancellation);
// 1 beyond (t < 2.1e-9 rounds away).  Within 2.2 ulp of tanh; odd, with both
// zeros mapped to +0.  IEEE single ops only, restated op for op by the oracle.
SCE_HD float det_tanhf(float z) {
  const float ax = __builtin_fabsf(z);
  float r;
  if (ax < 0.25f) {
    const float x2 = ax * ax;
    float p = -0x1.226e36p-7f;
    p = p * x2 + 0x1.664f48p-6f;
    p = p * x2 + -0x1.ba1ba2p-5f;
    p = p * x2 + 0x1.111112p-3f;
    p = p * x2 + -0x1.555556p-2f;
    r = ax + ax * (x2 * p);
  } else if (ax > 10.0f) {
    r = 1.0f;
  } else {
    const float y = -2.0f * ax;
    const float k = __builtin_rintf(y * 0x1.715476p+0f);
    const float q = (y - k * 0x1.63p-1f) - k * -0x1.bd0106p-13f;
    float p = 0x1.a01a02p-13f;
    p = p * q + 0x1.6c16c2p-10f;
    p = p * q + 0x1.111112p-7f;
    p = p * q + 0x1.555556p-5f;
    p = p * q + 0x1.555556p-3f;
    p = p * q + 0.5f;
    p = p * q + 1.0f;
    p = p * q + 1.0f;
    const float t = p * __builtin_bit_cast(float, (uint32_t)((int)k + 127) << 23);
    r = (1.0f - t) / (1.0f + t);
  }
  return z < 0.0f ? -r : r;
}

SCE_HD double det_log(double x) {
  if (x != x || x < 0.0) return __builtin_nan("");
  if (x == 0.0) return -__builtin_inf();
  if (x == __builtin_inf()) return x;
  int eadj = 0;
  if (x < 0x1p-1022) { x = x * 0x1p54; eadj = -54; }
  const uint64_t b = d_to_bits(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023 + eadj;
  double m = bits_to_d((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > 0x1.6a09e667f3bcdp+0) { m = m * 0.5; e += 1; }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double t = 0x1.47ae147ae147bp-5;
  t = t * z + 0x1.642c8590b2164p-5;
  t = t * z + 0x1.8618618618618p-5;
  t = t * z + 0x1.af286bca1af28p-5;
  t = t * z + 0x1.e1e1e1e1e1e1ep-5;
  t = t * z + 0x1.1111111111111p-4;
  t = t * z + 0x1.3b13b13b13b14p-4;
  t = t * z + 0x1.745d1745d1746p-4;
  t = t * z + 0x1.c71c71c71c71cp-4;
  t = t * z + 0x1.2492492492492p-3;
  t = t * z + 0x1.999999999999ap-3;
  t = t * z + 0x1.5555555555555p-2;
  const double R = t * z;
  const double s2 = s + s;
  const double lm = s2 + s2 * R;
  const double de = (double)e;
  return de * 0x1.62e42fee00000p-1 + (de * 0x1.a39ef35793c76p-33 + lm);
}

SCE_HD void det_sincos(double x, double* so, double* co) {
  const double n = __builtin_rint(x * 0x1.45f306dc9c883p-1);
  const double r = ((x - n * 0x1.921fb54400000p+0) - n * 0x1.0b4611a600000p-34) -
                   n * 0x1.3198a2e037073p-69;
  const double z = r * r;
  double ts = -0x1.2f49b46814157p-57;
  ts = ts * z + 0x1.952c77030ad4ap-49;
  ts = ts * z + -0x1.ae7f3e733b81fp-41;
  ts = ts * z + 0x1.6124613a86d09p-33;
  ts = ts * z + -0x1.ae64567f544e4p-26;
  ts = ts * z + 0x1.71de3a556c734p-19;
  ts = ts * z + -0x1.a01a01a01a01ap-13;
  ts = ts * z + 0x1.1111111111111p-7;
  ts = ts * z + -0x1.5555555555555p-3;
  const double sn = r + (r * z) * ts;
  double tc = 0x1.e542ba4020225p-62;
  tc = tc * z + -0x1.6827863b97d97p-53;
  tc = tc * z + 0x1.ae7f3e733b81fp-45;
  tc = tc * z + -0x1.93974a8c07c9dp-37;
  tc = tc * z + 0x1.1eed8eff8d898p-29;
  tc = tc * z + -0x1.27e4fb7789f5cp-22;
  tc = tc * z + 0x1.a01a01a01a01ap-16;
  tc = tc * z + -0x1.6c16c16c16c17p-10;
  tc = tc * z + 0x1.5555555555555p-5;
  tc = tc * z + -0x1.0000000000000p-1;
  const double cs = 1.0 + z * tc;
  const int q = (int)((long long)n & 3);
  if (q == 0) { *so = sn; *co = cs; }
  else if (q == 1) { *so = cs; *co = -sn; }
  else if (q == 2) { *so = -sn; *co = -cs; }
  else { *so = -cs; *co = sn; }
}

// Box–Muller pair of one block — philox.jl:100-105.
SCE_HD void normal2(u32x4 b, double* n0, double* n1) {
  const double u1 = uniform01(b.x, b.y);
  const double u2 = uniform01(b.z, b.w);
  const double r = __builtin_sqrt(-2.0 * det_log(u1));
  double s, c;
  det_sincos(6.283185307179586 * u2, &s, &c);
  *n0 = r * c;
  *n1 = r * s;
}

// ---------------------------------------------------------------------------
// Tesseral row Z_lm(u), l = 0..LMAX, in lm_index order — the operation order of
// src/gpu/zlm_device.jl:151-263 (LegendrePolynomials v0.4 dnPl, Harmonics
// _plm_norm / Zlm_unsafe, Base.power_by_squaring for (x+iy)^n).
// ---------------------------------------------------------------------------
SCE_HD double plm_norm(int l, int m) {
  double acc = (double)(2 * l + 1) / (4.0 * 3.141592653589793);
  for (int i = l - m + 1; i <= l + m; ++i) acc /= (double)i;
  return __builtin_sqrt(acc);
}

struct cplx { double re, im; };
SCE_HD cplx cmul(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

SCE_HD cplx cpow_int(cplx x, int p) {
  if (p == 0) return cplx{1.0, 0.0};
  const cplx xsq = cmul(x, x);
  if (p == 1) return x;
  if (p == 2) return xsq;
  int t = __builtin_ctz((unsigned)p) + 1;
  p >>= t;
  if ((t -= 1) > 0) x = xsq;
  while ((t -= 1) > 0) x = cmul(x, x);
  cplx y = x;
  while (p > 0) {
    t = __builtin_ctz((unsigned)p) + 1;
    p >>= t;
    while ((t -= 1) >= 0) x = cmul(x, x);
    y = cmul(y, x);
  }
  return y;
}

template <int LMAX>
SCE_HD double dnpl(double x, int l, int n, double* cache) {
  if (n == l) {
    double p = 1.0;
    for (int i = 2 * l - 1; i >= 1; i -= 2) p *= (double)i;
    cache[0] = p;
  } else {
    double Pl = 1.0, Plm1 = 0.0;
    cache[0] = Pl;
    for (int ll = 1; ll <= l - n; ++ll) {
      const double nP = ((double)(2 * ll - 1) * x * Pl - (double)(ll - 1) * Plm1) / (double)ll;
      Plm1 = Pl;
      Pl = nP;
      cache[ll] = Pl;
    }
    for (int ni = 1; ni <= n; ++ni) {
      const double pnn = (double)(2 * ni - 1) * cache[0];
      cache[0] = pnn;
      const int ll = ni + 1;
      cache[1] = ((double)(2 * ll - 1) * (x * pnn + (double)ni * cache[1])) / (double)ll;
      const int top = (l < l - n + ni) ? l : l - n + ni;
      for (int li = ni + 2; li <= top; ++li) {
        cache[li - ni] = ((double)(2 * li - 1) * (x * cache[li - ni - 1] +
                                                  (double)ni * cache[li - ni]) -
                          (double)(li - 1) * cache[li - ni - 2]) / (double)li;
      }
    }
  }
  return cache[l - n];
}

template <int LMAX>
SCE_HD void zlm_row(double ux, double uy, double uz, double* z) {
  double cache[LMAX + 2];
  int i = 0;
#pragma unroll
  for (int l = 0; l <= LMAX; ++l) {
#pragma unroll
    for (int m = -l; m <= l; ++m) {
      const int n = m < 0 ? -m : m;
      const double par = (n & 1) ? -1.0 : 1.0;
      const double plm = par * plm_norm(l, n) * dnpl<LMAX>(uz, l, n, cache);
      if (n == 0) {
        z[i++] = plm;
      } else {
        const double c = par * __builtin_sqrt(2.0) * plm;
        const cplx zp = cpow_int(cplx{ux, uy}, n);
        z[i++] = m > 0 ? c * zp.re : c * zp.im;
      }
    }
  }
}

// Runtime-lmax dispatch for host callers.
SCE_HD void zlm_row_dyn(int lmax, double ux, double uy, double uz, double* z) {
  switch (lmax) {
    case 0: zlm_row<0>(ux, uy, uz, z); break;
    case 1: zlm_row<1>(ux, uy, uz, z); break;
    case 2: zlm_row<2>(ux, uy, uz, z); break;
    case 3: zlm_row<3>(ux, uy, uz, z); break;
    case 4: zlm_row<4>(ux, uy, uz, z); break;
    case 5: zlm_row<5>(ux, uy, uz, z); break;
    default: zlm_row<6>(ux, uy, uz, z); break;
  }
}

// ---------------------------------------------------------------------------
// Tangent-projected gradient of one tesseral harmonic, grad Z_lm(u) — the
// operation order of src/gpu/grad_device.jl:30-66 (Harmonics.grad_Zlm_unsafe):
// the order-n and order-(n+1) dnPl at u_z with parity and normalisation (the
// order-(n+1) derivative is a +0.0 literal when n + 1 > l), the Cartesian
// derivative of c P^(n)(z) Re/Im (x + iy)^n, and the radial part
// u (u . dZ) / |u|^2 removed.  Only + - * / and a correctly rounded sqrt.
// ---------------------------------------------------------------------------
template <int LMAX>
SCE_HD void grad_zlm_one(int l, int m, double x, double y, double z, double g[3]) {
  double cache[LMAX + 2];
  const int n = m < 0 ? -m : m;
  const double par = (n & 1) ? -1.0 : 1.0;
  const double plm = par * plm_norm(l, n) * dnpl<LMAX>(z, l, n, cache);
  const double d1 = l < n + 1 ? 0.0 : dnpl<LMAX>(z, l, n + 1, cache);
  const double dplm = par * plm_norm(l, n) * d1;
  const double r2 = x * x + y * y + z * z;
  if (m == 0) {
    const double zz = z * dplm / r2;
    g[0] = -x * zz;
    g[1] = -y * zz;
    g[2] = dplm - z * zz;
    return;
  }
  const double c = par * __builtin_sqrt(2.0);
  const cplx pn = cpow_int(cplx{x, y}, n), pn1 = cpow_int(cplx{x, y}, n - 1);
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

// lm index k (lm order) -> (l, m)
SCE_HD void lm_of(int k, int* l, int* m) {
  int ll = 0;
  while ((ll + 1) * (ll + 1) <= k) ++ll;
  *l = ll;
  *m = k - ll * ll - ll;
}

// ---------------------------------------------------------------------------
// Keyed Metropolis proposal — src/gpu/gpu_sweep.jl:32-41 with
// _FLIP_FRACTION = 0.2 and _rotate (src/updates.jl:20,25-29).
// normalize(v) = v / sqrt((x*x + y*y) + z*z).
// ---------------------------------------------------------------------------
SCE_HD void rotate3(const double e[3], const double a[3], double th, double out[3]) {
  double s, c;
  det_sincos(th, &s, &c);
  const double cr0 = a[1] * e[2] - a[2] * e[1];
  const double cr1 = a[2] * e[0] - a[0] * e[2];
  const double cr2 = a[0] * e[1] - a[1] * e[0];
  const double d = (a[0] * e[0] + a[1] * e[1]) + a[2] * e[2];
  const double g = (1.0 - c) * d;
  out[0] = (c * e[0] + s * cr0) + g * a[0];
  out[1] = (c * e[1] + s * cr1) + g * a[1];
  out[2] = (c * e[2] + s * cr2) + g * a[2];
}

SCE_HD void keyed_proposal(uint64_t seed, int32_t site, int32_t sweep, const double e[3],
                           double step, double e2[3], double* u_acc) {
  const u32x4 b0 = philox_block(seed, site, sweep, 0u, 0u);
  const double u_flip = uniform01(b0.x, b0.y);
  *u_acc = uniform01(b0.z, b0.w);
  if (u_flip < 0.2) {
    e2[0] = -e[0]; e2[1] = -e[1]; e2[2] = -e[2];
    return;
  }
  double n1, n2, n3, n4;
  normal2(philox_block(seed, site, sweep, 1u, 0u), &n1, &n2);
  normal2(philox_block(seed, site, sweep, 2u, 0u), &n3, &n4);
  const double nn = __builtin_sqrt((n1 * n1 + n2 * n2) + n3 * n3);
  const double ax[3] = {n1 / nn, n2 / nn, n3 / nn};
  rotate3(e, ax, step * n4, e2);
}

// Philox slots of the library: 0..2 are the reference's Metropolis slots
// (philox.jl:65-67); 3 is the delayed-acceptance second-stage uniform; 16+r is
// the accept uniform of overrelaxation pass r of a compound sweep.
constexpr uint32_t SLOT_DA2 = 3u;

// 1 / N1, N1 = sqrt(3 / (4 pi)) the l = 1 tesseral normalisation, rounded to
// double: the surrogate descriptor reads a neighbour's spin back from its row as
// (Z[3], Z[1], Z[2]) * INV_N1.
constexpr double INV_N1 = 0x1.05f8bd37c0e62p+1;
constexpr uint32_t SLOT_OR0 = 16u;

}  // namespace sce
