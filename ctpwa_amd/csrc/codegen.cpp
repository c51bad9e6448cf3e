// Source generator for the fused per-event likelihood kernel.
//
// The kernel is specialised to one model: the number of SL tensors, polarisations,
// amplitudes, blocks, propagator denominators and free parameters are baked in, so
// every per-event quantity and every accumulator lives in registers and all index
// arithmetic is constant.  It is compiled once per model with hipRTC.
//
// Per event e (one thread, workgroup-contiguous chunks of one event class):
//   Psi_d  = 1 / D_d(e)                      each propagator denominator
//   Psi_b  = prod_d Psi_d                    each block
//   S_p    = sum_a c_a Psi_b(a) sl'_u(a),p   c_a = v_a K_a (theta-only)
//   I      = sum_p |S_p|^2
//   w_p    = S_p (phase space) | clamp(w_e S_p / I, +-1e10) (data, background;
//            src/ComputeNLL.cu:21-84), I == 0 -> log(1e-10), w = 0
//   T_u    = sum_p conj(sl'_u,p) w_p
//   accX_a += conj(Psi_b(a)) T_u(a)          -> d(ln-like)/dv*  (times K_a later)
//   R_d    = Z_b Psi_b Psi_d,  Z_b = sum_{a in b} c_a conj(T_u(a))
//   accT_t += x_t Re R_d | y_t Im R_d        one per theta-dependent denominator term;
//            dL/dtheta_j = -2 sum_t (dcoef_t/dtheta_j) accT_t is formed after the sum
//   accV   += I (phase space) | w_e ln I (data, background)
// Inputs are read straight from HBM, or (pipelined form) copied tile by tile into LDS
// by the DMA path while the previous tile is computed; a deterministic wave/LDS
// reduction then writes one partial row per workgroup.
#include <algorithm>
#include <functional>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "model.hpp"

namespace ctpwa {

static std::string D(double x);
static std::string DH_(double x) { return x < 0 ? "(-1.0)" : "1.0"; }
static std::string D(double x) {
    std::ostringstream s;
    s.precision(17);
    s << x;
    std::string r = s.str();
    if (r.find_first_of(".eEn") == std::string::npos) r += ".0";
    return r;
}


// Complex numbers carrying first derivatives over NCV block-local parameters (forward
// mode), for Custom lineshapes; the operations and functions follow the expression
// language's complex semantics (principal log, exp(b log a) powers, principal square
// root, real q0/Bf/BW/BWR model functions as in the built-in lineshapes).
static const char* CD_PREAMBLE = R"(
struct CD { double r, i; double dr[NCV], di[NCV]; };
__device__ __forceinline__ CD cd_c(double r, double i) {
  CD z; z.r = r; z.i = i;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = 0.0; z.di[k] = 0.0; }
  return z;
}
__device__ __forceinline__ CD cd_var(double v, int j) { CD z = cd_c(v, 0.0); z.dr[j] = 1.0; return z; }
__device__ __forceinline__ CD cd_add(const CD& a, const CD& b) {
  CD z; z.r = a.r + b.r; z.i = a.i + b.i;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = a.dr[k] + b.dr[k]; z.di[k] = a.di[k] + b.di[k]; }
  return z;
}
__device__ __forceinline__ CD cd_sub(const CD& a, const CD& b) {
  CD z; z.r = a.r - b.r; z.i = a.i - b.i;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = a.dr[k] - b.dr[k]; z.di[k] = a.di[k] - b.di[k]; }
  return z;
}
__device__ __forceinline__ CD cd_neg(const CD& a) {
  CD z; z.r = -a.r; z.i = -a.i;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = -a.dr[k]; z.di[k] = -a.di[k]; }
  return z;
}
__device__ __forceinline__ CD cd_mul(const CD& a, const CD& b) {
  CD z; z.r = a.r * b.r - a.i * b.i; z.i = a.r * b.i + a.i * b.r;
  for (int k = 0; k < NCV; ++k) {
    z.dr[k] = a.dr[k] * b.r - a.di[k] * b.i + a.r * b.dr[k] - a.i * b.di[k];
    z.di[k] = a.dr[k] * b.i + a.di[k] * b.r + a.r * b.di[k] + a.i * b.dr[k];
  }
  return z;
}
// value f(a) with f'(a) = (pr + i pi): derivative chain rule
__device__ __forceinline__ CD cd_chain(double fr, double fi, double pr, double pi, const CD& a) {
  CD z; z.r = fr; z.i = fi;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = pr * a.dr[k] - pi * a.di[k]; z.di[k] = pr * a.di[k] + pi * a.dr[k]; }
  return z;
}
__device__ __forceinline__ CD cd_inv(const CD& a) {
  const double q = 1.0 / (a.r * a.r + a.i * a.i);
  const double vr = a.r * q, vi = -a.i * q;                 // 1/a
  return cd_chain(vr, vi, -(vr * vr - vi * vi), -(2.0 * vr * vi), a);   // d(1/a) = -da/a^2
}
__device__ __forceinline__ CD cd_div(const CD& a, const CD& b) { return cd_mul(a, cd_inv(b)); }
__device__ __forceinline__ CD cd_exp(const CD& a) {
  const double e = exp(a.r), vr = e * cos(a.i), vi = e * sin(a.i);
  return cd_chain(vr, vi, vr, vi, a);
}
__device__ __forceinline__ CD cd_log(const CD& a) {
  const double q = 1.0 / (a.r * a.r + a.i * a.i);
  return cd_chain(0.5 * log(a.r * a.r + a.i * a.i), atan2(a.i, a.r), a.r * q, -a.i * q, a);
}
__device__ __forceinline__ CD cd_pow(const CD& a, const CD& b) { return cd_exp(cd_mul(b, cd_log(a))); }
__device__ __forceinline__ CD cd_sin(const CD& a) {
  const double ep = exp(a.i), em = exp(-a.i);
  const double sr = sin(a.r) * (ep + em) * 0.5, si = cos(a.r) * (ep - em) * 0.5;
  const double cr = cos(a.r) * (ep + em) * 0.5, ci = -sin(a.r) * (ep - em) * 0.5;
  return cd_chain(sr, si, cr, ci, a);
}
__device__ __forceinline__ CD cd_cos(const CD& a) {
  const double ep = exp(a.i), em = exp(-a.i);
  const double sr = sin(a.r) * (ep + em) * 0.5, si = cos(a.r) * (ep - em) * 0.5;
  const double cr = cos(a.r) * (ep + em) * 0.5, ci = -sin(a.r) * (ep - em) * 0.5;
  return cd_chain(cr, ci, -sr, -si, a);
}
__device__ __forceinline__ CD cd_sqrt(const CD& a) {
  double vr, vi;
  if (a.i == 0.0) {
    if (a.r >= 0.0) { vr = sqrt(a.r); vi = 0.0; } else { vr = 0.0; vi = sqrt(-a.r); }
  } else {
    const double r = sqrt(a.r * a.r + a.i * a.i);
    vr = sqrt((r + a.r) * 0.5);
    vi = sqrt((r - a.r) * 0.5);
    if (a.i < 0.0) vi = -vi;
  }
  // d sqrt(a) = da / (2 sqrt(a))
  const double q = 0.5 / (vr * vr + vi * vi);
  return cd_chain(vr, vi, vr * q, -vi * q, a);
}
__device__ __forceinline__ CD cd_abs(const CD& a) {
  const double v = sqrt(a.r * a.r + a.i * a.i);
  CD z = cd_c(v, 0.0);
  for (int k = 0; k < NCV; ++k) z.dr[k] = (a.r * a.dr[k] + a.i * a.di[k]) / v;
  return z;
}
__device__ __forceinline__ CD cd_re(const CD& a) {
  CD z = cd_c(a.r, 0.0);
  for (int k = 0; k < NCV; ++k) z.dr[k] = a.dr[k];
  return z;
}
__device__ __forceinline__ CD cd_im(const CD& a) {
  CD z = cd_c(a.i, 0.0);
  for (int k = 0; k < NCV; ++k) z.dr[k] = a.di[k];
  return z;
}
__device__ __forceinline__ CD cd_conj(const CD& a) {
  CD z; z.r = a.r; z.i = -a.i;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = a.dr[k]; z.di[k] = -a.di[k]; }
  return z;
}
// real parts only from here on: breakup at x (computeQ0AD), 0 below threshold
__device__ __forceinline__ CD cd_q0(const CD& x, const CD& m1, const CD& m2) {
  const double s = m1.r + m2.r, d = m1.r - m2.r, x2 = x.r * x.r;
  const double q2 = (x2 - s * s) * (x2 - d * d) / (4.0 * x2);
  if (q2 < 0.0) return cd_c(0.0, 0.0);
  const double q = sqrt(q2);
  // dq2/dx = (x^4 - s^2 d^2) / (2 x^3)
  const double dq2 = (x2 * x2 - s * s * d * d) / (2.0 * x2 * x.r);
  CD z = cd_c(q, 0.0);
  for (int k = 0; k < NCV; ++k) z.dr[k] = (q > 0.0) ? dq2 / (2.0 * q) * x.dr[k] : 0.0;
  return z;
}
__device__ __forceinline__ void bfpoly_d(int L, double z, double& n, double& dn) {
  // N_L(z) with z the argument q d, and dN/dz
  const double z2 = z * z;
  double p, dp;   // polynomial in z2 and its derivative
  switch (L) {
  case 1: p = 1.0 + z2; dp = 1.0; break;
  case 2: p = 9.0 + z2 * (3.0 + z2); dp = 3.0 + 2.0 * z2; break;
  case 3: p = 225.0 + z2 * (45.0 + z2 * (6.0 + z2)); dp = 45.0 + z2 * (12.0 + 3.0 * z2); break;
  case 4: p = 11025.0 + z2 * (1575.0 + z2 * (135.0 + z2 * (10.0 + z2)));
          dp = 1575.0 + z2 * (270.0 + z2 * (30.0 + 4.0 * z2)); break;
  case 5: p = 893025.0 + z2 * (99225.0 + z2 * (6300.0 + z2 * (315.0 + z2 * (15.0 + z2))));
          dp = 99225.0 + z2 * (12600.0 + z2 * (945.0 + z2 * (60.0 + 5.0 * z2))); break;
  case 6: p = 540326025.0 + z2 * (6185025.0 + z2 * (363825.0 + z2 * (17325.0 + z2 * (630.0 + z2 * (21.0 + z2)))));
          dp = 6185025.0 + z2 * (727650.0 + z2 * (51975.0 + z2 * (2520.0 + z2 * (105.0 + 6.0 * z2)))); break;
  default: p = 1.0; dp = 0.0;
  }
  n = p;
  dn = dp * 2.0 * z;
}
// Bf(L, q, q0, d) = sqrt(N_L(q0 d) / N_L(q d))
__device__ __forceinline__ CD cd_bf(int L, const CD& q, const CD& q0, const CD& d) {
  if (L == 0 || L > 6) return cd_c(1.0, 0.0);
  double n0, dn0, n1, dn1;
  bfpoly_d(L, q0.r * d.r, n0, dn0);
  bfpoly_d(L, q.r * d.r, n1, dn1);
  const double v = sqrt(n0 / n1);
  CD z = cd_c(v, 0.0);
  for (int k = 0; k < NCV; ++k) {
    const double dn0k = dn0 * (q0.dr[k] * d.r + q0.r * d.dr[k]);
    const double dn1k = dn1 * (q.dr[k] * d.r + q.r * d.dr[k]);
    z.dr[k] = 0.5 * v * (dn0k / n0 - dn1k / n1);
  }
  return z;
}
// Flatte channel phase-space factor of the event mass m (real part above threshold,
// imaginary part below); the channel masses are taken as constants
__device__ __forceinline__ double rho_val(double m, double ma, double mb, int im) {
  const double s = m * m;
  const double f1 = 1.0 - (ma + mb) * (ma + mb) / s, f2 = 1.0 - (ma - mb) * (ma - mb) / s;
  const double g2 = f2 > 0.0 ? sqrt(f2) : 0.0;
  return im ? ((f1 >= 0.0 ? 0.0 : sqrt(-f1)) * g2) : ((f1 >= 0.0 ? sqrt(f1) : 0.0) * g2);
}
__device__ __forceinline__ CD cd_rho(double m, const CD& ma, const CD& mb, int im) { return cd_c(rho_val(m, ma.r, mb.r, im), 0.0); }
// 1 / (m0^2 - m^2 - i m0 g0)
__device__ __forceinline__ CD cd_bw(const CD& m, const CD& m0, const CD& g0) {
  return cd_inv(cd_sub(cd_sub(cd_mul(cd_re(m0), cd_re(m0)), cd_mul(cd_re(m), cd_re(m))),
                       cd_mul(cd_c(0.0, 1.0), cd_mul(cd_re(m0), cd_re(g0)))));
}
// 1 / (m0^2 - m^2 - i m0 g0 (q/q0)^(2L+1) (m0/m) Bf^2)
__device__ __forceinline__ CD cd_bwr(int L, const CD& m, const CD& m0, const CD& g0, const CD& q, const CD& q0,
                                     const CD& d) {
  const CD bf = cd_bf(L, q, q0, d);
  CD r = cd_div(cd_re(q), cd_re(q0));
  CD p = cd_c(1.0, 0.0);
  for (int k = 0; k < 2 * L + 1; ++k) p = cd_mul(p, r);
  const CD gam = cd_mul(cd_mul(cd_mul(cd_re(g0), p), cd_div(cd_re(m0), cd_re(m))), cd_mul(bf, bf));
  return cd_inv(cd_sub(cd_sub(cd_mul(cd_re(m0), cd_re(m0)), cd_mul(cd_re(m), cd_re(m))),
                       cd_mul(cd_c(0.0, 1.0), cd_mul(cd_re(m0), gam))));
}
)";

// Complex numbers carrying first and second derivatives over NCV block-local parameters
// (second-order forward mode) for the Hessian of models with Custom lineshapes.  Second
// derivatives are packed over k <= l (NH = NCV (NCV + 1) / 2 entries).  Every function
// f applies z_k = f'(a) a_k, z_kl = f'(a) a_kl + f''(a) a_k a_l with the same branches as
// the first-order CD functions.
static const char* CH_PREAMBLE = R"(
#define NH (NCV * (NCV + 1) / 2)
__device__ __forceinline__ int hx(int k, int l) { return k * NCV - (k * (k - 1)) / 2 + (l - k); }
struct CH { double r, i; double dr[NCV], di[NCV]; double hr[NH], hi[NH]; };
__device__ __forceinline__ CH ch_c(double r, double i) {
  CH z; z.r = r; z.i = i;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = 0.0; z.di[k] = 0.0; }
  for (int h = 0; h < NH; ++h) { z.hr[h] = 0.0; z.hi[h] = 0.0; }
  return z;
}
__device__ __forceinline__ CH ch_var(double v, int j) { CH z = ch_c(v, 0.0); z.dr[j] = 1.0; return z; }
__device__ __forceinline__ CH ch_add(const CH& a, const CH& b) {
  CH z; z.r = a.r + b.r; z.i = a.i + b.i;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = a.dr[k] + b.dr[k]; z.di[k] = a.di[k] + b.di[k]; }
  for (int h = 0; h < NH; ++h) { z.hr[h] = a.hr[h] + b.hr[h]; z.hi[h] = a.hi[h] + b.hi[h]; }
  return z;
}
__device__ __forceinline__ CH ch_sub(const CH& a, const CH& b) {
  CH z; z.r = a.r - b.r; z.i = a.i - b.i;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = a.dr[k] - b.dr[k]; z.di[k] = a.di[k] - b.di[k]; }
  for (int h = 0; h < NH; ++h) { z.hr[h] = a.hr[h] - b.hr[h]; z.hi[h] = a.hi[h] - b.hi[h]; }
  return z;
}
__device__ __forceinline__ CH ch_neg(const CH& a) {
  CH z; z.r = -a.r; z.i = -a.i;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = -a.dr[k]; z.di[k] = -a.di[k]; }
  for (int h = 0; h < NH; ++h) { z.hr[h] = -a.hr[h]; z.hi[h] = -a.hi[h]; }
  return z;
}
// (ar + i ai) (br + i bi)
#define CMR(ar, ai, br, bi) ((ar) * (br) - (ai) * (bi))
#define CMI(ar, ai, br, bi) ((ar) * (bi) + (ai) * (br))
__device__ __forceinline__ CH ch_mul(const CH& a, const CH& b) {
  CH z; z.r = CMR(a.r, a.i, b.r, b.i); z.i = CMI(a.r, a.i, b.r, b.i);
  for (int k = 0; k < NCV; ++k) {
    z.dr[k] = CMR(a.dr[k], a.di[k], b.r, b.i) + CMR(a.r, a.i, b.dr[k], b.di[k]);
    z.di[k] = CMI(a.dr[k], a.di[k], b.r, b.i) + CMI(a.r, a.i, b.dr[k], b.di[k]);
  }
  for (int k = 0; k < NCV; ++k)
    for (int l = k; l < NCV; ++l) {
      const int h = hx(k, l);
      z.hr[h] = CMR(a.hr[h], a.hi[h], b.r, b.i) + CMR(a.dr[k], a.di[k], b.dr[l], b.di[l]) +
                CMR(a.dr[l], a.di[l], b.dr[k], b.di[k]) + CMR(a.r, a.i, b.hr[h], b.hi[h]);
      z.hi[h] = CMI(a.hr[h], a.hi[h], b.r, b.i) + CMI(a.dr[k], a.di[k], b.dr[l], b.di[l]) +
                CMI(a.dr[l], a.di[l], b.dr[k], b.di[k]) + CMI(a.r, a.i, b.hr[h], b.hi[h]);
    }
  return z;
}
// value f(a) with f'(a) = p, f''(a) = q (complex)
__device__ __forceinline__ CH ch_chain(double fr, double fi, double pr, double pi, double qr, double qi, const CH& a) {
  CH z; z.r = fr; z.i = fi;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = CMR(pr, pi, a.dr[k], a.di[k]); z.di[k] = CMI(pr, pi, a.dr[k], a.di[k]); }
  for (int k = 0; k < NCV; ++k)
    for (int l = k; l < NCV; ++l) {
      const int h = hx(k, l);
      const double ur = CMR(a.dr[k], a.di[k], a.dr[l], a.di[l]), ui = CMI(a.dr[k], a.di[k], a.dr[l], a.di[l]);
      z.hr[h] = CMR(pr, pi, a.hr[h], a.hi[h]) + CMR(qr, qi, ur, ui);
      z.hi[h] = CMI(pr, pi, a.hr[h], a.hi[h]) + CMI(qr, qi, ur, ui);
    }
  return z;
}
__device__ __forceinline__ CH ch_inv(const CH& a) {
  const double q = 1.0 / (a.r * a.r + a.i * a.i);
  const double vr = a.r * q, vi = -a.i * q;                   // 1/a
  const double v2r = vr * vr - vi * vi, v2i = 2.0 * vr * vi;  // 1/a^2
  return ch_chain(vr, vi, -v2r, -v2i, 2.0 * CMR(v2r, v2i, vr, vi), 2.0 * CMI(v2r, v2i, vr, vi), a);
}
__device__ __forceinline__ CH ch_div(const CH& a, const CH& b) { return ch_mul(a, ch_inv(b)); }
__device__ __forceinline__ CH ch_exp(const CH& a) {
  const double e = exp(a.r), vr = e * cos(a.i), vi = e * sin(a.i);
  return ch_chain(vr, vi, vr, vi, vr, vi, a);
}
__device__ __forceinline__ CH ch_log(const CH& a) {
  const double q = 1.0 / (a.r * a.r + a.i * a.i);
  const double ur = a.r * q, ui = -a.i * q;                   // 1/a
  return ch_chain(0.5 * log(a.r * a.r + a.i * a.i), atan2(a.i, a.r), ur, ui, -(ur * ur - ui * ui), -2.0 * ur * ui, a);
}
__device__ __forceinline__ CH ch_pow(const CH& a, const CH& b) { return ch_exp(ch_mul(b, ch_log(a))); }
__device__ __forceinline__ CH ch_sin(const CH& a) {
  const double ep = exp(a.i), em = exp(-a.i);
  const double sr = sin(a.r) * (ep + em) * 0.5, si = cos(a.r) * (ep - em) * 0.5;
  const double cr = cos(a.r) * (ep + em) * 0.5, ci = -sin(a.r) * (ep - em) * 0.5;
  return ch_chain(sr, si, cr, ci, -sr, -si, a);
}
__device__ __forceinline__ CH ch_cos(const CH& a) {
  const double ep = exp(a.i), em = exp(-a.i);
  const double sr = sin(a.r) * (ep + em) * 0.5, si = cos(a.r) * (ep - em) * 0.5;
  const double cr = cos(a.r) * (ep + em) * 0.5, ci = -sin(a.r) * (ep - em) * 0.5;
  return ch_chain(cr, ci, -sr, -si, -cr, -ci, a);
}
__device__ __forceinline__ CH ch_sqrt(const CH& a) {
  double vr, vi;
  if (a.i == 0.0) {
    if (a.r >= 0.0) { vr = sqrt(a.r); vi = 0.0; } else { vr = 0.0; vi = sqrt(-a.r); }
  } else {
    const double r = sqrt(a.r * a.r + a.i * a.i);
    vr = sqrt((r + a.r) * 0.5);
    vi = sqrt((r - a.r) * 0.5);
    if (a.i < 0.0) vi = -vi;
  }
  // f' = 1 / (2 sqrt a), f'' = -1 / (4 sqrt(a)^3)
  const double q = 1.0 / (vr * vr + vi * vi);
  const double ur = vr * q, ui = -vi * q;                     // 1/sqrt(a)
  const double u2r = ur * ur - ui * ui, u2i = 2.0 * ur * ui;
  return ch_chain(vr, vi, 0.5 * ur, 0.5 * ui, -0.25 * CMR(u2r, u2i, ur, ui), -0.25 * CMI(u2r, u2i, ur, ui), a);
}
__device__ __forceinline__ CH ch_abs(const CH& a) {
  const double v = sqrt(a.r * a.r + a.i * a.i);
  CH z = ch_c(v, 0.0);
  for (int k = 0; k < NCV; ++k) z.dr[k] = (a.r * a.dr[k] + a.i * a.di[k]) / v;
  for (int k = 0; k < NCV; ++k)
    for (int l = k; l < NCV; ++l) {
      const int h = hx(k, l);
      z.hr[h] = (a.dr[k] * a.dr[l] + a.di[k] * a.di[l] + a.r * a.hr[h] + a.i * a.hi[h] - z.dr[k] * z.dr[l]) / v;
    }
  return z;
}
__device__ __forceinline__ CH ch_re(const CH& a) {
  CH z = ch_c(a.r, 0.0);
  for (int k = 0; k < NCV; ++k) z.dr[k] = a.dr[k];
  for (int h = 0; h < NH; ++h) z.hr[h] = a.hr[h];
  return z;
}
__device__ __forceinline__ CH ch_im(const CH& a) {
  CH z = ch_c(a.i, 0.0);
  for (int k = 0; k < NCV; ++k) z.dr[k] = a.di[k];
  for (int h = 0; h < NH; ++h) z.hr[h] = a.hi[h];
  return z;
}
__device__ __forceinline__ CH ch_conj(const CH& a) {
  CH z; z.r = a.r; z.i = -a.i;
  for (int k = 0; k < NCV; ++k) { z.dr[k] = a.dr[k]; z.di[k] = -a.di[k]; }
  for (int h = 0; h < NH; ++h) { z.hr[h] = a.hr[h]; z.hi[h] = -a.hi[h]; }
  return z;
}
// real function g(x) of the real part of a with g' = p, g'' = q
__device__ __forceinline__ CH ch_real_chain(double g, double p, double q, const CH& a) {
  return ch_chain(g, 0.0, p, 0.0, q, 0.0, ch_re(a));
}
// breakup at x with constant daughter masses (computeQ0AD), 0 below threshold
__device__ __forceinline__ CH ch_q0(const CH& x, const CH& m1, const CH& m2) {
  const double s = m1.r + m2.r, d = m1.r - m2.r, x2 = x.r * x.r;
  const double q2 = (x2 - s * s) * (x2 - d * d) / (4.0 * x2);
  if (q2 < 0.0) return ch_c(0.0, 0.0);
  const double q = sqrt(q2);
  if (q == 0.0) return ch_c(0.0, 0.0);
  const double dq2 = (x2 * x2 - s * s * d * d) / (2.0 * x2 * x.r);          // dq2/dx
  const double d2q2 = 0.5 + 1.5 * s * s * d * d / (x2 * x2);                  // d2q2/dx2
  return ch_real_chain(q, dq2 / (2.0 * q), d2q2 / (2.0 * q) - dq2 * dq2 / (4.0 * q * q2), x);
}
__device__ __forceinline__ void bfpoly_d2(int L, double z, double& n, double& dn, double& d2n) {
  // N_L(z) = P(z^2): N' = 2 z P'(u), N'' = 2 P'(u) + 4 z^2 P''(u), u = z^2
  const double u = z * z;
  double p, dp, d2p;
  switch (L) {
  case 1: p = 1.0 + u; dp = 1.0; d2p = 0.0; break;
  case 2: p = 9.0 + u * (3.0 + u); dp = 3.0 + 2.0 * u; d2p = 2.0; break;
  case 3: p = 225.0 + u * (45.0 + u * (6.0 + u)); dp = 45.0 + u * (12.0 + 3.0 * u); d2p = 12.0 + 6.0 * u; break;
  case 4: p = 11025.0 + u * (1575.0 + u * (135.0 + u * (10.0 + u)));
          dp = 1575.0 + u * (270.0 + u * (30.0 + 4.0 * u)); d2p = 270.0 + u * (60.0 + 12.0 * u); break;
  case 5: p = 893025.0 + u * (99225.0 + u * (6300.0 + u * (315.0 + u * (15.0 + u))));
          dp = 99225.0 + u * (12600.0 + u * (945.0 + u * (60.0 + 5.0 * u)));
          d2p = 12600.0 + u * (1890.0 + u * (180.0 + 20.0 * u)); break;
  case 6: p = 540326025.0 + u * (6185025.0 + u * (363825.0 + u * (17325.0 + u * (630.0 + u * (21.0 + u)))));
          dp = 6185025.0 + u * (727650.0 + u * (51975.0 + u * (2520.0 + u * (105.0 + 6.0 * u))));
          d2p = 727650.0 + u * (103950.0 + u * (7560.0 + u * (420.0 + 30.0 * u))); break;
  default: p = 1.0; dp = 0.0; d2p = 0.0;
  }
  n = p;
  dn = 2.0 * z * dp;
  d2n = 2.0 * dp + 4.0 * u * d2p;
}
// Bf(L, q, q0, d) = sqrt(N_L(q0 d) / N_L(q d)) = exp(g), g = (ln N(z0) - ln N(z1)) / 2
__device__ __forceinline__ CH ch_bf(int L, const CH& q, const CH& q0, const CH& d) {
  if (L == 0 || L > 6) return ch_c(1.0, 0.0);
  double n0, dn0, d2n0, n1, dn1, d2n1;
  bfpoly_d2(L, q0.r * d.r, n0, dn0, d2n0);
  bfpoly_d2(L, q.r * d.r, n1, dn1, d2n1);
  const double v = sqrt(n0 / n1);
  const double a0 = dn0 / n0, b0 = d2n0 / n0 - a0 * a0, a1 = dn1 / n1, b1 = d2n1 / n1 - a1 * a1;
  double z0[NCV], z1[NCV], g[NCV];
  for (int k = 0; k < NCV; ++k) {
    z0[k] = q0.dr[k] * d.r + q0.r * d.dr[k];
    z1[k] = q.dr[k] * d.r + q.r * d.dr[k];
    g[k] = 0.5 * (a0 * z0[k] - a1 * z1[k]);
  }
  CH z = ch_c(v, 0.0);
  for (int k = 0; k < NCV; ++k) z.dr[k] = v * g[k];
  for (int k = 0; k < NCV; ++k)
    for (int l = k; l < NCV; ++l) {
      const int h = hx(k, l);
      const double z0kl = q0.hr[h] * d.r + q0.dr[k] * d.dr[l] + q0.dr[l] * d.dr[k] + q0.r * d.hr[h];
      const double z1kl = q.hr[h] * d.r + q.dr[k] * d.dr[l] + q.dr[l] * d.dr[k] + q.r * d.hr[h];
      const double gkl = 0.5 * (b0 * z0[k] * z0[l] + a0 * z0kl - b1 * z1[k] * z1[l] - a1 * z1kl);
      z.hr[h] = v * (gkl + g[k] * g[l]);
    }
  return z;
}
__device__ __forceinline__ double rho_val(double m, double ma, double mb, int im) {
  const double s = m * m;
  const double f1 = 1.0 - (ma + mb) * (ma + mb) / s, f2 = 1.0 - (ma - mb) * (ma - mb) / s;
  const double g2 = f2 > 0.0 ? sqrt(f2) : 0.0;
  return im ? ((f1 >= 0.0 ? 0.0 : sqrt(-f1)) * g2) : ((f1 >= 0.0 ? sqrt(f1) : 0.0) * g2);
}
__device__ __forceinline__ CH ch_rho(double m, const CH& ma, const CH& mb, int im) { return ch_c(rho_val(m, ma.r, mb.r, im), 0.0); }
__device__ __forceinline__ CH ch_bw(const CH& m, const CH& m0, const CH& g0) {
  return ch_inv(ch_sub(ch_sub(ch_mul(ch_re(m0), ch_re(m0)), ch_mul(ch_re(m), ch_re(m))),
                       ch_mul(ch_c(0.0, 1.0), ch_mul(ch_re(m0), ch_re(g0)))));
}
__device__ __forceinline__ CH ch_bwr(int L, const CH& m, const CH& m0, const CH& g0, const CH& q, const CH& q0,
                                     const CH& d) {
  const CH bf = ch_bf(L, q, q0, d);
  CH r = ch_div(ch_re(q), ch_re(q0));
  CH p = ch_c(1.0, 0.0);
  for (int k = 0; k < 2 * L + 1; ++k) p = ch_mul(p, r);
  const CH gam = ch_mul(ch_mul(ch_mul(ch_re(g0), p), ch_div(ch_re(m0), ch_re(m))), ch_mul(bf, bf));
  return ch_inv(ch_sub(ch_sub(ch_mul(ch_re(m0), ch_re(m0)), ch_mul(ch_re(m), ch_re(m))),
                       ch_mul(ch_c(0.0, 1.0), ch_mul(ch_re(m0), gam))));
}
)";

// code of one Custom expression (prefix tokens) as a nested CD expression
// (prefix "cd_": first-order duals; "ch_": second-order duals)
static std::string custom_code(const std::vector<std::string>& t, size_t& i, const CustomSpec& C,
                               const std::vector<int>& vars, const std::string& X = "cd_") {
    if (i >= t.size()) throw std::runtime_error("Custom expression is truncated");
    const std::string k = t[i++];
    auto sub = [&]() { return custom_code(t, i, C, vars, X); };
    auto num = [](const std::string& x) {
        std::ostringstream s;
        s.precision(17);
        s << std::stod(x);
        std::string r = s.str();
        if (r.find_first_of(".eEn") == std::string::npos) r += ".0";
        return r;
    };
    if (k.rfind("n:", 0) == 0) return X + "c(" + num(k.substr(2)) + ", 0.0)";
    if (k == "I") return X + "c(0.0, 1.0)";
    if (k == "v:m") return X + "c(xm, 0.0)";
    if (k == "v:q") return X + "c(xq, 0.0)";
    {
        static const char* dvar[8] = {"v:p1p", "v:p1e", "v:p1costheta", "v:p1phi",
                                      "v:p2p", "v:p2e", "v:p2costheta", "v:p2phi"};
        for (int j = 0; j < 8; ++j)
            if (k == dvar[j]) return X + "c(xd" + std::to_string(j) + ", 0.0)";
    }
    if (k == "rr" || k == "ri") {
        std::string a = sub(), b = sub();
        return X + "rho(xm, " + a + ", " + b + (k == "ri" ? ", 1)" : ", 0)");
    }
    if (k.rfind("p:", 0) == 0) {
        int pk = std::stoi(k.substr(2));
        if (pk < 0 || pk >= (int)C.param_val.size()) throw std::runtime_error("Custom parameter index out of range");
        int s = C.param_slot[pk];
        if (s < 0) return X + "c(" + num(std::to_string(C.param_val[pk])) + ", 0.0)";
        for (size_t li = 0; li < vars.size(); ++li)
            if (vars[li] == s) return X + "var(tb[OFFT + " + std::to_string(s) + "], " + std::to_string(li) + ")";
        throw std::runtime_error("Custom parameter slot is not a block variable");
    }
    if (k == "+") { std::string a = sub(), b = sub(); return X + "add(" + a + ", " + b + ")"; }
    if (k == "-") { std::string a = sub(), b = sub(); return X + "sub(" + a + ", " + b + ")"; }
    if (k == "*") { std::string a = sub(), b = sub(); return X + "mul(" + a + ", " + b + ")"; }
    if (k == "/") { std::string a = sub(), b = sub(); return X + "div(" + a + ", " + b + ")"; }
    if (k == "^") { std::string a = sub(), b = sub(); return X + "pow(" + a + ", " + b + ")"; }
    if (k == "neg") return X + "neg(" + sub() + ")";
    if (k.rfind("f:", 0) == 0) {
        std::string f = k.substr(2);
        static const char* ok[] = {"exp", "log", "sin", "cos", "sqrt", "abs", "re", "im", "conj"};
        bool known = false;
        for (const char* o : ok) if (f == o) known = true;
        if (!known) throw std::runtime_error("Custom expression: unknown function " + f);
        return X + f + "(" + sub() + ")";
    }
    if (k == "q0") { std::string a = sub(), b = sub(), c = sub(); return X + "q0(" + a + ", " + b + ", " + c + ")"; }
    if (k.rfind("bf:", 0) == 0) {
        std::string L = std::to_string(std::stoi(k.substr(3)));
        std::string a = sub(), b = sub(), c = sub();
        return X + "bf(" + L + ", " + a + ", " + b + ", " + c + ")";
    }
    if (k == "bw") { std::string a = sub(), b = sub(), c = sub(); return X + "bw(" + a + ", " + b + ", " + c + ")"; }
    if (k.rfind("bwr:", 0) == 0) {
        std::string L = std::to_string(std::stoi(k.substr(4)));
        std::string a = sub(), b = sub(), c = sub(), d = sub(), e = sub(), f = sub();
        return X + "bwr(" + L + ", " + a + ", " + b + ", " + c + ", " + d + ", " + e + ", " + f + ")";
    }
    throw std::runtime_error("Custom expression: unknown token " + k);
}

// The second-order Custom factors for the Hessian, in a module of their own (compiled on
// the first Hessian call): rows [c * NCF2 + (r, i, dr[NCV], di[NCV], hr[NH], hi[NH])]
// of the output for every stored event.
std::string generate_custom2_kernel(const Model_& M) {
    if (M.customs.empty()) throw std::runtime_error("model has no Custom lineshapes");
    std::ostringstream o;
    const int ncv = M.ncv(), nh = ncv * (ncv + 1) / 2, ncf2 = 2 + 2 * ncv + 2 * nh;
    o << "// generated: second-order Custom lineshape factors\n"
      << "#define NCV " << ncv << "\n#define OFFT " << M.off_th() << "LL\n" << CH_PREAMBLE;
    std::vector<int> imap = M.imap();
    auto inv_load = [&](size_t i) {
        std::ostringstream q;
        if (imap[i] >= 0) q << "inv[" << imap[i] << "LL * ld + e]";
        else q << "(double)invf[" << (-imap[i] - 1) << "LL * ld + e]";
        return q.str();
    };
    o << "extern \"C\" __global__ void __launch_bounds__(64) ctpwa_custom2(\n"
         "    const double* __restrict__ tab, const double* __restrict__ inv, const float* __restrict__ invf,\n"
         "    long long ld, double* __restrict__ out)\n{\n"
         "  const double* tb = tab;\n"
         "  for (long long e = (long long)blockIdx.x * 64 + threadIdx.x; e < ld; e += (long long)gridDim.x * 64) {\n";
    for (size_t c = 0; c < M.customs.size(); ++c) {
        const CustomSpec& C = M.customs[c];
        size_t ti = 0;
        std::string code = custom_code(C.expr, ti, C, M.bvars[C.block], "ch_");
        if (ti != C.expr.size()) throw std::runtime_error("Custom expression has trailing tokens");
        const long long r0 = (long long)c * ncf2;
        o << "    {\n      const double xm = " << inv_load(C.inv_m) << ", xq = " << inv_load(C.inv_q) << ";\n";
        for (int j = 0; j < 8; ++j)
            if (C.inv_d[j] >= 0) o << "      const double xd" << j << " = " << inv_load(C.inv_d[j]) << ";\n";
        o
          << "      const CH F = " << code << ";\n"
          << "      out[" << r0 << "LL * ld + e] = F.r; out[" << r0 + 1 << "LL * ld + e] = F.i;\n"
          << "      for (int k = 0; k < NCV; ++k) { out[(" << r0 + 2 << "LL + k) * ld + e] = F.dr[k]; out[(" << r0 + 2
          << "LL + NCV + k) * ld + e] = F.di[k]; }\n"
          << "      for (int h = 0; h < NH; ++h) { out[(" << r0 + 2 + 2 * ncv << "LL + h) * ld + e] = F.hr[h]; out[("
          << r0 + 2 + 2 * ncv << "LL + NH + h) * ld + e] = F.hi[h]; }\n    }\n";
    }
    o << "  }\n}\n";
    return o.str();
}

// Denominator coefficients that the prep kernel computes by the same operations from the
// same inputs, so that their values are equal bit for bit: m0^2 (BWR, BW, Flatte) of the
// same mass source (theta slot, or fixed value), the BW width term m0 g0 of the same mass
// and width, and the BWR width term g0 m0^2 N_L(q0 d) / q0^(2L+1) of the same mass, width,
// L, barrier radius and fixed daughter masses (the rho states of the pi+- pi0 chains of a
// three-pion model).  The fused kernel loads each distinct one once: fewer scalar
// registers, which otherwise run out and spill to vector-register lanes that every event
// iteration reads back.  alias[d][c] = (d' << 8) | c' of the first equal coefficient, or -1.
static std::vector<std::vector<int>> shared_coefficients(const Model_& M) {
    const Plan& P = M.P;
    std::vector<std::vector<int>> alias(M.dens.size());
    std::map<std::string, int> first;
    auto src = [](const PRes& R, int k) {
        std::ostringstream q;
        q.precision(17);
        if ((int)R.params.size() <= k) q << "none";
        else if ((int)R.slots.size() > k && R.slots[k] >= 0) q << "slot" << R.slots[k];
        else q << "val" << R.params[k];
        return q.str();
    };
    for (size_t d = 0; d < M.dens.size(); ++d) {
        const DenSpec& D = M.dens[d];
        alias[d].assign(D.ncoef, -1);
        if (D.stat || D.ncoef > 255) continue;
        const PRes& R = P.blocks[D.block].res[D.res];
        const PNode& n = P.topos[P.blocks[D.block].topo].nodes[R.node];
        if (R.model != M_BWR && R.model != M_BW && R.model != M_FLATTE) continue;
        for (int c = 0; c < D.ncoef && c < 2; ++c) {
            std::ostringstream key;
            key.precision(17);
            if (c == 0) {
                key << "m2|" << (R.params.empty() ? std::string("one") : src(R, 0));
            } else if (R.model == M_BW) {
                key << "bw1|" << src(R, 0) << "|" << src(R, 1);
            } else if (R.model == M_BWR) {
                // q0 from fixed daughter masses only (a cascade's daughter mass parameter
                // makes q0 depend on the block)
                if (!(n.md1 > 0 || n.d1ref >= 0) || !(n.md2 > 0 || n.d2ref >= 0)) continue;
                key << "bwr1|" << src(R, 0) << "|" << src(R, 1) << "|" << n.lmin << "|" << (n.has_bf ? n.bf_d : 0.0) << "|"
                    << (n.md1 > 0 ? n.md1 : 0.0) << "|" << (n.md2 > 0 ? n.md2 : 0.0);
            } else {
                continue;
            }
            auto it = first.find(key.str());
            if (it == first.end()) first[key.str()] = ((int)d << 8) | c;
            else alias[d][c] = it->second;
        }
    }
    return alias;
}

std::string generate_eval_kernel(const Model_& M) {
    const Plan& P = M.P;
    int NU = (int)M.us.size(), NP = P.npol, NA = P.namp;
    int NTH = (int)M.thterms.size();
    int NACC = M.nacc();
    std::vector<int> used(M.inv.size(), 0);
    for (const DenSpec& Dn : M.dens) {
        for (const Term& t : Dn.re) if (t.inv >= 0) used[t.inv] = 1;
        for (const Term& t : Dn.im) if (t.inv >= 0) used[t.inv] = 1;
    }
    std::ostringstream o;
    long long tile_bytes = 16;       // LDS tile buffer of the pipelined form (per buffer)
    o << "// generated: fused likelihood kernel\n";
    if (!M.customs.empty())
        o << "#define NCV " << M.ncv() << "\n#define NCF " << M.ncf() << "\n#define OFFT " << M.off_th() << "LL\n"
          << CD_PREAMBLE;
    // reciprocal: hardware estimate refined by two Newton steps (~1 ulp)
    o << "__device__ __forceinline__ double frcp(double x) {\n"
         "  double r = __builtin_amdgcn_rcp(x);\n"
         "  double e = __builtin_fma(-x, r, 1.0); r = __builtin_fma(r, e, r);\n"
         "  e = __builtin_fma(-x, r, 1.0); return __builtin_fma(r, e, r);\n}\n";
    o << "#define OFFC " << M.off_c() << "LL\n#define OFFR " << M.off_rec() << "LL\n#define REC " << Model_::REC << "\n";
    // A wave-uniform double read as two 32-bit halves through readfirstlane: the value is
    // known to be uniform and lives in scalar registers.  A load through a pointer that is
    // a kernel argument becomes a scalar load by itself; one through a pointer read from
    // memory (a job descriptor) does not, and the 40-odd coefficients then took vector
    // registers and halved the grouped kernel's occupancy.
    o << "__device__ __forceinline__ double uld(const double* p) {\n"
         "  const unsigned long long b = *(const unsigned long long*)p;\n"
         "  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);\n"
         "  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));\n"
         "  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);\n}\n";
    // one workgroup's share of the events: the body is shared by the single-likelihood
    // kernel and the grouped kernel that evaluates many likelihoods of this model in one
    // launch, so a likelihood's sums do not depend on how it was launched (UL: how the
    // coefficients are read, which does not change their values)
    o << "template <bool UL>\n__device__ __forceinline__ void eval_body(\n"
         "    const double* __restrict__ tab, const void* __restrict__ slv, const double* __restrict__ inv,\n"
         "    const float* __restrict__ invf, const double* __restrict__ wts, double* __restrict__ part,\n"
         "    const int* __restrict__ wg_cls, const long long* __restrict__ wg_rng, long long ld, int nwg,\n"
         "    double w1, double w2, int pw1, int pw2, const int wg, double (*__restrict__ red)[" << NACC << "],\n"
         "    unsigned char* __restrict__ lbuf, const double* __restrict__ cfb)\n{\n";
    // storage types: FP64, or FP32 SL tensors / FP32 invariants except s (mixed)
    const char* SLT = M.mixed ? "float2" : "double2";
    o << "  const " << SLT << "* __restrict__ sl = (const " << SLT << "*)slv;\n";
    std::vector<int> imap = M.imap();
    auto inv_load = [&](size_t i, const std::string& ev) {
        std::ostringstream q;
        if (imap[i] >= 0) q << "inv[" << imap[i] << "LL * ld + " << ev << "]";
        else q << "(double)invf[" << (-imap[i] - 1) << "LL * ld + " << ev << "]";
        return q.str();
    };
    auto sl_conv = [&](const std::string& x) {
        return M.mixed ? "make_double2((double)" + x + ".x, (double)" + x + ".y)" : x;
    };
    // stored polarisations per SL tensor; a polarisation past them is the signed
    // conjugate of its mirror image
    const int NPS = M.npstore();
    auto mirror = [&](int u, int p, const char* ind) {
        std::ostringstream q;
        const int m = NP - 1 - p;
        const double g = (double)M.psign[u];
        q << ind << "const double2 s" << u << "_" << p << " = make_double2(" << DH_(g) << " * s" << u << "_" << m << ".x, "
          << DH_(-g) << " * s" << u << "_" << m << ".y);\n";
        return q.str();
    };
    o << "  const int cls = wg_cls[wg];\n"
         "  const long long e0 = wg_rng[2 * wg], e1 = wg_rng[2 * wg + 1];\n"
         "  const double wc = (cls == 1) ? w1 : w2;\n  const int pw = (cls == 1) ? pw1 : ((cls == 2) ? pw2 : 0);\n"
         "  const double LOG_FLOOR = log(1e-10);\n";
    // theta-only coefficients (wave-uniform)
    o << "  const double* tb = tab;\n";
    const std::vector<std::vector<int>> coef_alias = shared_coefficients(M);
    // ul: inside eval_body, whose UL instantiation (the grouped kernels, where the table
    // pointer comes from a job descriptor in memory rather than a kernel argument) reads
    // the coefficients through uld() so that they stay in scalar registers
    auto emit_coefs = [&](const char* ind, bool ul) {
        auto ld = [&](const std::string& x) {
            return ul ? "(UL ? uld(tb + (" + x + ")) : tb[" + x + "])" : "tb[" + x + "]";
        };
        for (int a = 0; a < NA; ++a)
            o << ind << "const double c" << a << "r = " << ld("OFFC + " + std::to_string(2 * a)) << ", c" << a
              << "i = " << ld("OFFC + " + std::to_string(2 * a + 1)) << ";\n";
        for (size_t d = 0; d < M.dens.size(); ++d)
            for (int c = 0; c < M.dens[d].ncoef; ++c) {
                const int src = coef_alias[d][c];
                if (src >= 0) {
                    o << ind << "const double k" << d << "_" << c << " = k" << (src >> 8) << "_" << (src & 255) << ";\n";
                    continue;
                }
                o << ind << "const double k" << d << "_" << c << " = "
                  << ld("OFFR + " + std::to_string((long long)(M.dens[d].coef_base + c)) + "LL * REC") << ";\n";
            }
    };
    emit_coefs("  ", true);
    o << "  double av = 0.0;\n";
    // constant-weight data and background classes: sum_e w ln I_e = w ln prod_e I_e, the
    // product kept as a mantissa in [0.5, 1) and a binary exponent (two frexp per event
    // instead of a logarithm; one logarithm per thread after the loop)
    o << "  double lm = 1.0;\n  int le = 0;\n";
    for (int a = 0; a < NA; ++a) o << "  double ax" << a << "r = 0.0, ax" << a << "i = 0.0;\n";
    for (int t = 0; t < NTH; ++t) o << "  double at" << t << " = 0.0;\n";
    const int NCT = (int)M.cterms.size();
    for (int t = 0; t < NCT; ++t) o << "  double ct" << t << " = 0.0;\n";
    const int WGS = M.wgs, NW = M.wgs / 64;
    if (M.lds_pipe) {
        // Tiles of WGS events: every per-event input row of the next tile is copied
        // HBM -> LDS by the DMA path (global_load_lds_dwordx4, no registers held) while
        // the current tile is computed from LDS.  Row r of a tile occupies
        // [roff_r, roff_r + WGS * es_r) of the tile buffer; the event layout (class
        // offsets and leading dimension multiples of 64 events) keeps every 16-byte chunk
        // aligned and inside its row.
        struct Row { std::string src; int es; long long off; std::string rd; };
        std::vector<Row> rows;
        long long tb = 0;
        const int esl = M.mixed ? 8 : 16;
        for (int u = 0; u < NU; ++u)
            for (int p = 0; p < NP; ++p) {
                std::string nm = std::to_string(u) + "_" + std::to_string(p);
                if (p >= NPS) {
                    rows.push_back(Row{std::string(), 0, tb, mirror(u, p, "    ")});
                    continue;
                }
                std::ostringstream src;
                src << "(const unsigned char*)sl + " << (long long)u * NPS + p << "LL * ld * " << esl;
                std::ostringstream rd;
                rd << "    const " << SLT << " g" << nm << " = ((const " << SLT << "*)(B + " << tb << "))[threadIdx.x];\n"
                   << "    const double2 s" << nm << " = " << sl_conv("g" + nm) << ";\n";
                rows.push_back(Row{src.str(), esl, tb, rd.str()});
                tb += (long long)WGS * esl;
            }
        for (size_t i = 0; i < M.inv.size(); ++i) {
            if (!used[i]) continue;
            std::ostringstream src, rd;
            if (imap[i] >= 0) {
                src << "(const unsigned char*)inv + " << imap[i] << "LL * ld * 8";
                rd << "    const double x" << i << " = ((const double*)(B + " << tb << "))[threadIdx.x];\n";
                rows.push_back(Row{src.str(), 8, tb, rd.str()});
                tb += (long long)WGS * 8;
            } else {
                src << "(const unsigned char*)invf + " << (-imap[i] - 1) << "LL * ld * 4";
                rd << "    const double x" << i << " = (double)((const float*)(B + " << tb << "))[threadIdx.x];\n";
                rows.push_back(Row{src.str(), 4, tb, rd.str()});
                tb += (long long)WGS * 4;
            }
        }
        o << "  typedef __attribute__((address_space(3))) void* lds_ptr;\n";
        o << "  const int ntile = (int)((e1 - e0 + " << WGS - 1 << ") / " << WGS << ");\n";
        // issue the copies of the tile starting at event ts into buffer dst
        o << "  auto issue = [&](unsigned char* dst, const long long ts) {\n"
             "    const int t = threadIdx.x;\n";
        for (const Row& r : rows) {
            if (r.es == 0) continue;             // a mirrored polarisation: no copy
            int nl = WGS * r.es / 16;
            o << "    if (t < " << nl << " && ts + (long long)(t * " << 16 / r.es << ") < e1)\n"
              << "      __builtin_amdgcn_global_load_lds((const void*)(" << r.src << " + ts * " << r.es << " + t * 16),\n"
              << "                                     (lds_ptr)(dst + " << r.off << " + (t & ~63) * 16), 16, 0, 0);\n";
        }
        o << "  };\n";
        o << "  issue(lbuf, e0);\n";
        o << "  for (int it = 0; it < ntile; ++it) {\n"
             "    const long long ts = e0 + (long long)it * " << WGS << ";\n"
             "    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this thread's copies landed\n"
             "    __syncthreads();                      // every thread's copies landed; previous tile consumed\n"
             "    if (it + 1 < ntile) issue(lbuf + ((it + 1) & 1) * " << tb << ", ts + " << WGS << ");\n"
             "    const unsigned char* B = lbuf + (it & 1) * " << tb << ";\n"
             "    const long long e = ts + threadIdx.x;\n"
             "    if (e < e1) {\n";
        for (const Row& r : rows) o << r.rd;
        tile_bytes = tb;
    } else {
        if (M.prefetch_sl) {
            // the SL values and invariants of the thread's next event (depth 1) or next two
            // events (depth 2: register sets n and m, rotated each iteration) are loaded
            // ahead in registers, in their stored types, so their latency overlaps the
            // current event's arithmetic
            const int depth = M.prefetch_sl;
            auto inv_raw = [&](size_t i, const std::string& ev) {
                std::ostringstream q;
                if (imap[i] >= 0) q << "inv[" << imap[i] << "LL * ld + " << ev << "]";
                else q << "invf[" << (-imap[i] - 1) << "LL * ld + " << ev << "]";
                return q.str();
            };
            auto decl_set = [&](const char* pre, const char* ev) {
                for (int u = 0; u < NU; ++u)
                    for (int p = 0; p < NPS; ++p)
                        o << "  " << SLT << " " << pre << u << "_" << p << " = sl[" << (long long)u * NPS + p << "LL * ld + " << ev
                          << "];\n";
                for (size_t i = 0; i < M.inv.size(); ++i)
                    if (used[i]) o << "  " << (imap[i] >= 0 ? "double" : "float") << " " << pre << "x" << i << " = " << inv_raw(i, ev) << ";\n";
            };
            // the event weight of a class with per-event weights is loaded ahead too: a load
            // issued after the next-event loads would make its wait cover them as well
            o << "  long long e = e0 + threadIdx.x;\n  const long long ep = (e < e1) ? e : e0;\n";
            decl_set("n", "ep");
            o << "  double nwt = pw ? wts[ep] : wc;\n";
            if (depth > 1) {
                o << "  const long long ep2 = (e + " << WGS << " < e1) ? e + " << WGS << " : e0;\n";
                decl_set("m", "ep2");
                o << "  double mwt = pw ? wts[ep2] : wc;\n";
            }
            o << "  for (; e < e1; e += " << WGS << ") {\n";
            for (int u = 0; u < NU; ++u)
                for (int p = 0; p < NP; ++p) {
                    std::string nm = std::to_string(u) + "_" + std::to_string(p);
                    if (p >= NPS) o << mirror(u, p, "    ");
                    else o << "    const double2 s" << nm << " = " << sl_conv("n" + nm) << ";\n";
                }
            for (size_t i = 0; i < M.inv.size(); ++i)
                if (used[i]) o << "    const double x" << i << " = (double)nx" << i << ";\n";
            o << "    const double wev = nwt;\n";
            const long long ahead = (long long)depth * WGS;
            o << "    { const long long en = (e + " << ahead << " < e1) ? e + " << ahead << " : e;\n";
            const char* tgt = depth > 1 ? "m" : "n";
            for (int u = 0; u < NU; ++u)
                for (int p = 0; p < NPS; ++p) {
                    if (depth > 1) o << "      n" << u << "_" << p << " = m" << u << "_" << p << ";\n";
                    o << "      " << tgt << u << "_" << p << " = sl[" << (long long)u * NPS + p << "LL * ld + en];\n";
                }
            for (size_t i = 0; i < M.inv.size(); ++i)
                if (used[i]) {
                    if (depth > 1) o << "      nx" << i << " = mx" << i << ";\n";
                    o << "      " << tgt << "x" << i << " = " << inv_raw(i, "en") << ";\n";
                }
            if (depth > 1) o << "      nwt = mwt;\n";
            o << "      if (pw) " << tgt << "wt = wts[en];\n";
            o << "    }\n";
        } else {
            o << "  for (long long e = e0 + threadIdx.x; e < e1; e += " << WGS << ") {\n";
            for (int u = 0; u < NU; ++u)
                for (int p = 0; p < NP; ++p) {
                    std::string nm = std::to_string(u) + "_" + std::to_string(p);
                    if (p >= NPS) {
                        o << mirror(u, p, "    ");
                        continue;
                    }
                    o << "    const " << SLT << " g" << nm << " = sl[" << (long long)u * NPS + p << "LL * ld + e];\n";
                    o << "    const double2 s" << nm << " = " << sl_conv("g" + nm) << ";\n";
                }
            for (size_t i = 0; i < M.inv.size(); ++i)
                if (used[i]) o << "    const double x" << i << " = " << inv_load(i, "e") << ";\n";
        }
    }
    // per-event core: denominators, Custom factors, block products (shared by the fused
    // kernel and the amplitude/intensity kernel of models with Custom lineshapes)
    auto emit_core = [&](std::ostringstream& o) {
        auto term_expr = [&](int d, const Term& t) {
            std::string coef = (t.coef < 0) ? D(t.cval) : ("k" + std::to_string(d) + "_" + std::to_string(t.coef));
            if (t.inv < 0) return coef;
            return coef + " * x" + std::to_string(t.inv);
        };
        // a sum starts from its first term (0.0 + x is an instruction: it is not x for x = -0.0)
        auto term_sum = [&](const char* nm, int d, const std::vector<Term>& ts) {
            if (ts.empty()) { o << "    double " << nm << d << " = 0.0;\n"; return; }
            o << "    double " << nm << d << " = " << term_expr(d, ts[0]) << ";\n";
            for (size_t k = 1; k < ts.size(); ++k) o << "    " << nm << d << " += " << term_expr(d, ts[k]) << ";\n";
        };
        // 1 / |D_d|^2 for groups of up to 8 propagators from one reciprocal: prefix products
        // a_k = m_0 ... m_k of m_d = |D_d|^2, r = 1 / a_last, then backwards q_k = r a_(k-1),
        // r *= m_k (Montgomery's simultaneous inversion).  One v_rcp_f64 with its Newton
        // steps per group instead of per propagator; each q_d is within a few ulp of 1/m_d, and
        // a group of 8 squared moduli stays far inside the FP64 range.
        std::vector<int> recip;
        for (size_t d = 0; d < M.dens.size(); ++d) {
            const DenSpec& Dn = M.dens[d];
            term_sum("dr", (int)d, Dn.re);
            term_sum("di", (int)d, Dn.im);
            if (Dn.stat) continue;   // table lineshape: the factor itself
            recip.push_back((int)d);
            o << "    const double m" << d << " = dr" << d << " * dr" << d << " + di" << d << " * di" << d << ";\n";
        }
        for (size_t g0 = 0; g0 < recip.size(); g0 += 8) {
            const size_t g1 = std::min(recip.size(), g0 + 8);
            if (g1 - g0 == 1) {
                const int d = recip[g0];
                o << "    const double q" << d << " = frcp(m" << d << ");\n";
                continue;
            }
            if (M.inv_tree) {
                // balanced product tree: the dependent chain is log2 of the group size deep
                // instead of the group size (same number of multiplications)
                auto nm = [&](size_t lo, size_t hi) {
                    return hi - lo == 1 ? "m" + std::to_string(recip[lo])
                                        : "u" + std::to_string(lo) + "_" + std::to_string(hi);
                };
                std::function<void(size_t, size_t)> prod = [&](size_t lo, size_t hi) {
                    if (hi - lo == 1) return;
                    const size_t mid = (lo + hi) / 2;
                    prod(lo, mid);
                    prod(mid, hi);
                    o << "    const double " << nm(lo, hi) << " = " << nm(lo, mid) << " * " << nm(mid, hi) << ";\n";
                };
                std::function<void(size_t, size_t, const std::string&)> inv = [&](size_t lo, size_t hi, const std::string& iv) {
                    if (hi - lo == 1) {
                        o << "    const double q" << recip[lo] << " = " << iv << ";\n";
                        return;
                    }
                    const size_t mid = (lo + hi) / 2;
                    const std::string il = "v" + std::to_string(lo) + "_" + std::to_string(mid);
                    const std::string ir = "v" + std::to_string(mid) + "_" + std::to_string(hi);
                    o << "    const double " << il << " = " << iv << " * " << nm(mid, hi) << ";\n";
                    o << "    const double " << ir << " = " << iv << " * " << nm(lo, mid) << ";\n";
                    inv(lo, mid, il);
                    inv(mid, hi, ir);
                };
                prod(g0, g1);
                const std::string rt = "v" + std::to_string(g0) + "_" + std::to_string(g1);
                o << "    const double " << rt << " = frcp(" << nm(g0, g1) << ");\n";
                inv(g0, g1, rt);
                continue;
            }
            for (size_t k = g0 + 1; k < g1; ++k) {
                o << "    const double a" << recip[k] << " = " << (k == g0 + 1 ? "m" : "a") << recip[k - 1] << " * m"
                  << recip[k] << ";\n";
            }
            o << "    double r" << recip[g0] << " = frcp(a" << recip[g1 - 1] << ");\n";
            for (size_t k = g1 - 1; k > g0; --k) {
                const int d = recip[k], prev = recip[k - 1];
                o << "    const double q" << d << " = r" << recip[g0] << " * " << (k - 1 == g0 ? "m" : "a") << prev << "; r"
                  << recip[g0] << " *= m" << d << ";\n";
            }
            o << "    const double q" << recip[g0] << " = r" << recip[g0] << ";\n";
        }
        for (size_t d = 0; d < M.dens.size(); ++d) {
            if (M.dens[d].stat) {
                o << "    const double pr" << d << " = dr" << d << ", pi" << d << " = di" << d << ";\n";
                continue;
            }
            o << "    const double pr" << d << " = dr" << d << " * q" << d << ", pi" << d << " = di" << d << " * q" << d << ";\n";
        }
        // Custom lineshape factors as written for this parameter point by ctpwa_custom (their
        // derivatives are loaded where the gradient accumulators use them)
        for (size_t c = 0; c < M.customs.size(); ++c) {
            const long long r0 = (long long)c * M.ncf();
            o << "    const double F" << c << "r = cfb[" << r0 << "LL * ld + e], F" << c << "i = cfb[" << r0 + 1
              << "LL * ld + e];\n";
        }
        for (size_t b = 0; b < P.blocks.size(); ++b) {
            const std::vector<int>& ds = M.block_dens[b];
            if (P.blocks[b].nsl == 0) continue;
            if (ds.empty()) o << "    double br" << b << " = 1.0, bi" << b << " = 0.0;\n";
            else o << "    double br" << b << " = pr" << ds[0] << ", bi" << b << " = pi" << ds[0] << ";\n";
            for (size_t k = 1; k < ds.size(); ++k) {
                int d = ds[k];
                o << "    { const double tr = br" << b << " * pr" << d << " - bi" << b << " * pi" << d
                  << "; bi" << b << " = br" << b << " * pi" << d << " + bi" << b << " * pr" << d << "; br" << b << " = tr; }\n";
            }
            if (!M.block_customs[b].empty()) {
                // product of the denominators alone (for the derivatives through Custom factors)
                o << "    const double pdr" << b << " = br" << b << ", pdi" << b << " = bi" << b << ";\n";
                for (int c : M.block_customs[b])
                    o << "    { const double tr = br" << b << " * F" << c << "r - bi" << b << " * F" << c << "i; bi" << b
                      << " = br" << b << " * F" << c << "i + bi" << b << " * F" << c << "r; br" << b << " = tr; }\n";
            }
        }
    };
    emit_core(o);
    for (int u = 0; u < NU; ++u) o << "    double y" << u << "r = 0.0, y" << u << "i = 0.0;\n";
    for (int a = 0; a < NA; ++a) {
        int b = M.amps[a].block, u = M.amps[a].u;
        o << "    y" << u << "r += c" << a << "r * br" << b << "; y" << u << "r -= c" << a << "i * bi" << b << "; y" << u
          << "i += c" << a << "r * bi" << b << "; y" << u << "i += c" << a << "i * br" << b << ";\n";
    }
    // mirror-image polarisations (one stored, sl_u,1 = g_u conj(sl_u,0)): with U_G = sum over
    // the SL tensors of sign g_u = G of Re(y_u) sl_u,0 and V_G the same with Im(y_u),
    // S_0 = U + i V and S_1 = conj(U') + i conj(V') (U = U_+ + U_-, U' = U_+ - U_-): 4
    // products per SL tensor for both polarisations instead of 8
    const bool mirror_pair = (NP == 2 && NPS == 1);
    bool gpos = false, gneg = false;
    for (int u = 0; u < NU && mirror_pair; ++u) (M.psign[u] > 0 ? gpos : gneg) = true;
    if (mirror_pair) {
        for (int G = 0; G < 2; ++G) {
            if (!(G == 0 ? gpos : gneg)) continue;
            const char* g = G == 0 ? "P" : "N";
            o << "    double U" << g << "r = 0.0, U" << g << "i = 0.0, V" << g << "r = 0.0, V" << g << "i = 0.0;\n";
            for (int u = 0; u < NU; ++u) {
                if ((M.psign[u] > 0) != (G == 0)) continue;
                o << "    U" << g << "r += y" << u << "r * s" << u << "_0.x; U" << g << "i += y" << u << "r * s" << u << "_0.y; V" << g
                  << "r += y" << u << "i * s" << u << "_0.x; V" << g << "i += y" << u << "i * s" << u << "_0.y;\n";
            }
        }
        if (gpos && gneg)
            o << "    const double Ur = UPr + UNr, Ui = UPi + UNi, Vr = VPr + VNr, Vi = VPi + VNi;\n"
                 "    const double Ur_ = UPr - UNr, Ui_ = UPi - UNi, Vr_ = VPr - VNr, Vi_ = VPi - VNi;\n";
        else if (gpos)
            o << "    const double Ur = UPr, Ui = UPi, Vr = VPr, Vi = VPi;\n"
                 "    const double Ur_ = UPr, Ui_ = UPi, Vr_ = VPr, Vi_ = VPi;\n";
        else
            o << "    const double Ur = UNr, Ui = UNi, Vr = VNr, Vi = VNi;\n"
                 "    const double Ur_ = -UNr, Ui_ = -UNi, Vr_ = -VNr, Vi_ = -VNi;\n";
        o << "    const double S0r = Ur - Vi, S0i = Ui + Vr;\n"
             "    const double S1r = Ur_ + Vi_, S1i = Vr_ - Ui_;\n";
    } else {
        for (int p = 0; p < NP; ++p) {
            o << "    double S" << p << "r = 0.0, S" << p << "i = 0.0;\n";
            for (int u = 0; u < NU; ++u)
                o << "    S" << p << "r += y" << u << "r * s" << u << "_" << p << ".x; S" << p << "r -= y" << u << "i * s" << u << "_" << p
                  << ".y; S" << p << "i += y" << u << "r * s" << u << "_" << p << ".y; S" << p << "i += y" << u << "i * s" << u << "_" << p << ".x;\n";
        }
    }
    o << "    double I = 0.0;\n";
    for (int p = 0; p < NP; ++p) o << "    I += S" << p << "r * S" << p << "r; I += S" << p << "i * S" << p << "i;\n";
    for (int p = 0; p < NP; ++p) o << "    double w" << p << "r, w" << p << "i;\n";
    o << "    if (cls == 0) {\n      av += I;\n";
    for (int p = 0; p < NP; ++p) o << "      w" << p << "r = S" << p << "r; w" << p << "i = S" << p << "i;\n";
    o << "    } else {\n      const double we = " << (M.prefetch_sl && !M.lds_pipe ? "wev" : "pw ? wts[e] : wc") << ";\n"
         "      if (pw) {\n        av += (I == 0.0 ? LOG_FLOOR : log(I)) * we;\n      } else {\n"
         "        const double Il = (I == 0.0) ? 1e-10 : I;   // ln(1e-10): the I = 0 rule\n"
         "        lm *= __builtin_amdgcn_frexp_mant(Il); le += __builtin_amdgcn_frexp_exp(Il);\n"
         "        const int ex = __builtin_amdgcn_frexp_exp(lm); lm = __builtin_amdgcn_frexp_mant(lm); le += ex;\n"
         "      }\n"
         "      if (I == 0.0) {\n";
    for (int p = 0; p < NP; ++p) o << "        w" << p << "r = 0.0; w" << p << "i = 0.0;\n";
    o << "      } else {\n        const double f = we * frcp(I);\n";
    for (int p = 0; p < NP; ++p)
        o << "        w" << p << "r = fmin(fmax(S" << p << "r * f, -1e10), 1e10); w" << p << "i = fmin(fmax(S" << p
          << "i * f, -1e10), 1e10);\n";
    o << "      }\n    }\n";
    if (mirror_pair) {
        // T_u = conj(sl_u,0) w_0 + g_u sl_u,0 w_1 from the sums and differences of w_0, w_1:
        // g = +1: (a sr + b di) + i (a si - b dr); g = -1: (a dr + b si) + i (a di - b sr)
        o << "    const double wsr = w0r + w1r, wsi = w0i + w1i, wdr = w0r - w1r, wdi = w0i - w1i;\n";
        for (int u = 0; u < NU; ++u) {
            const bool pos = M.psign[u] > 0;
            o << "    const double t" << u << "r = s" << u << "_0.x * " << (pos ? "wsr" : "wdr") << " + s" << u << "_0.y * "
              << (pos ? "wdi" : "wsi") << ", t" << u << "i = s" << u << "_0.x * " << (pos ? "wsi" : "wdi") << " - s" << u
              << "_0.y * " << (pos ? "wdr" : "wsr") << ";\n";
        }
    } else {
        for (int u = 0; u < NU; ++u) {
            o << "    double t" << u << "r = 0.0, t" << u << "i = 0.0;\n";
            for (int p = 0; p < NP; ++p)
                o << "    t" << u << "r += s" << u << "_" << p << ".x * w" << p << "r; t" << u << "r += s" << u << "_" << p << ".y * w" << p
                  << "i; t" << u << "i += s" << u << "_" << p << ".x * w" << p << "i; t" << u << "i -= s" << u << "_" << p << ".y * w" << p << "r;\n";
        }
    }
    // per block: the coupling sums P_a = conj(Psi_b) T_u (X_a += P_a), then the theta
    // terms R_d = Z_b Psi_b Psi_d with Z_b Psi_b = sum_{a in b} c_a conj(P_a)
    std::vector<std::vector<int>> den_terms(M.dens.size());
    for (int t = 0; t < NTH; ++t) den_terms[M.thterms[t].den].push_back(t);
    std::vector<std::vector<int>> custom_terms(M.customs.size());
    for (int t = 0; t < NCT; ++t) custom_terms[M.cterms[t].custom].push_back(t);
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const PBlock& B = P.blocks[b];
        for (int c = 0; c < B.nsl; ++c) {
            int a = B.site + c, u = M.amps[a].u;
            o << "    const double P" << a << "r = br" << b << " * t" << u << "r + bi" << b << " * t" << u << "i, P" << a
              << "i = br" << b << " * t" << u << "i - bi" << b << " * t" << u << "r;\n";
            o << "    ax" << a << "r += P" << a << "r; ax" << a << "i += P" << a << "i;\n";
        }
        bool any = false;
        for (int d : M.block_dens[b]) if (!den_terms[d].empty()) any = true;
        for (int c : M.block_customs[b]) if (!custom_terms[c].empty()) any = true;
        if (B.nsl == 0 || !any) continue;
        o << "    {\n";
        const bool with_custom = !M.block_customs[b].empty();
        if (with_custom) {
            // the derivatives through Custom factors need Z_b itself
            o << "      double zr = 0.0, zi = 0.0;\n";
            for (int c = 0; c < B.nsl; ++c) {
                int a = B.site + c, u = M.amps[a].u;
                o << "      zr += c" << a << "r * t" << u << "r; zr += c" << a << "i * t" << u << "i; zi += c" << a << "i * t" << u
                  << "r; zi -= c" << a << "r * t" << u << "i;\n";
            }
            o << "      const double zbr = zr * br" << b << " - zi * bi" << b << ", zbi = zr * bi" << b << " + zi * br" << b << ";\n";
        } else {
            o << "      double zbr = 0.0, zbi = 0.0;\n";
            for (int c = 0; c < B.nsl; ++c) {
                int a = B.site + c;
                o << "      zbr += c" << a << "r * P" << a << "r; zbr += c" << a << "i * P" << a << "i; zbi += c" << a
                  << "i * P" << a << "r; zbi -= c" << a << "r * P" << a << "i;\n";
            }
        }
        for (int d : M.block_dens[b]) {
            if (den_terms[d].empty()) continue;
            // a component with one term and no invariant is accumulated by two FMAs
            int nre = 0, nim = 0;
            bool bare_re = false, bare_im = false;
            for (int t : den_terms[d]) {
                const ThTerm& T = M.thterms[t];
                (T.im ? nim : nre)++;
                if (T.inv < 0) (T.im ? bare_im : bare_re) = true;
            }
            const bool fre = nre == 1 && bare_re, fim = nim == 1 && bare_im;
            o << "      {\n";
            if (nre && !fre) o << "        const double rr = zbr * pr" << d << " - zbi * pi" << d << ";\n";
            if (nim && !fim) o << "        const double ri = zbr * pi" << d << " + zbi * pr" << d << ";\n";
            for (int t : den_terms[d]) {
                const ThTerm& T = M.thterms[t];
                if (!T.im && fre) {
                    o << "        at" << t << " += zbr * pr" << d << "; at" << t << " -= zbi * pi" << d << ";\n";
                } else if (T.im && fim) {
                    o << "        at" << t << " += zbr * pi" << d << "; at" << t << " += zbi * pr" << d << ";\n";
                } else {
                    const char* comp = T.im ? "ri" : "rr";
                    if (T.inv < 0) o << "        at" << t << " += " << comp << ";\n";
                    else o << "        at" << t << " += x" << T.inv << " * " << comp << ";\n";
                }
            }
            o << "      }\n";
        }
        // Custom factors: d Psi_b / d theta_j = (product of the block's other factors) F'_c,j
        for (int c : M.block_customs[b]) {
            if (custom_terms[c].empty()) continue;
            o << "      {\n        double or_ = pdr" << b << ", oi_ = pdi" << b << ";\n";
            for (int c2 : M.block_customs[b]) {
                if (c2 == c) continue;
                o << "        { const double tr = or_ * F" << c2 << "r - oi_ * F" << c2 << "i; oi_ = or_ * F" << c2
                  << "i + oi_ * F" << c2 << "r; or_ = tr; }\n";
            }
            o << "        const double wr = zr * or_ - zi * oi_, wi = zr * oi_ + zi * or_;\n";
            for (int t : custom_terms[c]) {
                int li = M.cterms[t].local;
                const long long rr = (long long)c * M.ncf() + 2 + li, ri = rr + M.ncv();
                o << "        ct" << t << " += wr * cfb[" << rr << "LL * ld + e] - wi * cfb[" << ri << "LL * ld + e];\n";
            }
            o << "      }\n";
        }
        o << "    }\n";
    }
    o << (M.lds_pipe ? "    }\n  }\n" : "  }\n");
    // ln prod I = ln(mantissa) + exponent ln 2 (ln 2 split into a high part with trailing
    // zero bits and the remainder, so that the exponent's product stays exact)
    o << "  if (cls != 0 && !pw)\n"
         "    av += wc * __builtin_fma((double)le, 1.90821492927058770002e-10,\n"
         "                              __builtin_fma((double)le, 6.93147180369123816490e-01, log(lm)));\n";
    // reduction over the wave: 16 accumulators at a time are written to the wave's LDS slab
    // (rows padded to 17 doubles), lane 4k + q adds column k over lanes 16q .. 16q + 15 in
    // lane order, and two exchanges join the quarters (a fixed order); per 16 accumulators
    // that is 16 LDS writes, 16 reads and 2 exchanges instead of 6 exchanges each
    std::vector<std::string> accs;
    accs.push_back("av");
    for (int a = 0; a < NA; ++a) { accs.push_back("ax" + std::to_string(a) + "r"); accs.push_back("ax" + std::to_string(a) + "i"); }
    for (int t = 0; t < NTH; ++t) accs.push_back("at" + std::to_string(t));
    for (int t = 0; t < NCT; ++t) accs.push_back("ct" + std::to_string(t));
    if ((int)accs.size() != NACC) throw std::runtime_error("fused kernel generator: accumulator count mismatch");
    if (M.lds_pipe) o << "  __syncthreads();                 // every wave is done with the tile buffers\n";
    o << "  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n"
         "  {\n    double (*tl)[17] = (double (*)[17])(lbuf + wave * " << 64 * 17 * 8 << ");\n"
         "    const int k4 = lane >> 2, q4 = lane & 3;\n";
    for (int c0 = 0; c0 < NACC; c0 += 16) {
        const int m = std::min(16, NACC - c0);
        for (int k = 0; k < m; ++k) o << "    tl[lane][" << k << "] = " << accs[c0 + k] << ";\n";
        o << "    __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"wavefront\"); __builtin_amdgcn_wave_barrier();\n"
             "    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"wavefront\");\n"
             "    { double s = tl[16 * q4][k4];\n"
             "      for (int t = 1; t < 16; ++t) s += tl[16 * q4 + t][k4];\n"
             "      s += __shfl_xor(s, 1, 64); s += __shfl_xor(s, 2, 64);\n"
             "      if (q4 == 0 && k4 < " << m << ") red[wave][" << c0 << " + k4] = s; }\n"
             "    __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"wavefront\"); __builtin_amdgcn_wave_barrier();\n"
             "    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"wavefront\");\n";
    }
    o << "  }\n  __syncthreads();\n";
    // fixed-order combination of the wave partials
    std::string comb;
    if (NW == 1) comb = "red[0][k]";
    else if (NW == 2) comb = "red[0][k] + red[1][k]";
    else comb = "(red[0][k] + red[1][k]) + (red[2][k] + red[3][k])";
    o << "  for (int k = threadIdx.x; k < " << NACC << "; k += " << WGS << ")\n"
         "    part[(long long)k * nwg + wg] = " << comb << ";\n";
    o << "}\n";
    std::string lb = "__launch_bounds__(" + std::to_string(WGS) + (M.min_waves > 0 ? ", " + std::to_string(M.min_waves) : std::string()) + ")";
    std::ostringstream decl;
    decl << "  __shared__ double red[" << NW << "][" << NACC << "];\n"
         << "  __shared__ __attribute__((aligned(16))) unsigned char lbuf["
         << std::max<long long>(M.lds_pipe ? 2 * tile_bytes : 0, (long long)NW * 64 * 17 * 8) << "];\n";
    o << "extern \"C\" __global__ void " << lb << " ctpwa_eval(\n"
         "    const double* __restrict__ tab, const void* __restrict__ slv, const double* __restrict__ inv,\n"
         "    const float* __restrict__ invf, const double* __restrict__ wts, double* __restrict__ part,\n"
         "    const int* __restrict__ wg_cls, const long long* __restrict__ wg_rng, long long ld, int nwg,\n"
         "    double w1, double w2, int pw1, int pw2, const double* __restrict__ cfb)\n{\n" << decl.str() <<
         "  eval_body<false>(tab, slv, inv, invf, wts, part, wg_cls, wg_rng, ld, nwg, w1, w2, pw1, pw2, blockIdx.x, red, lbuf, cfb);\n}\n";
    // grouped launch: jobs[j] owns workgroups [wg_begin_j, wg_begin_{j+1}); the layout of
    // this struct is mirrored by EvalJob in capi.hip
    o << "struct EvalJob {\n"
         "  const double* tab; const void* sl; const double* inv; const float* invf; const double* wts;\n"
         "  double* part; const int* wg_cls; const long long* wg_rng; const double* cf; long long ld; double w1, w2;\n"
         "  int nwg, pw1, pw2, wg_begin;\n};\n";
    o << "extern \"C\" __global__ void " << lb << " ctpwa_eval_group(const EvalJob* __restrict__ jobs, int njobs)\n{\n"
      << decl.str() <<
         "  int lo = 0, hi = njobs - 1;\n"
         "  while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (jobs[mid].wg_begin <= (int)blockIdx.x) lo = mid; else hi = mid - 1; }\n"
         "  const EvalJob J = jobs[lo];\n"
         "  eval_body<true>(J.tab, J.sl, J.inv, J.invf, J.wts, J.part, J.wg_cls, J.wg_rng, J.ld, J.nwg, J.w1, J.w2, J.pw1, J.pw2,\n"
         "            (int)blockIdx.x - J.wg_begin, red, lbuf, J.cf);\n}\n";
    // persistent form (graph replays): hdr = {job count, total workgroup tasks}; each
    // workgroup takes tasks from the counter ctr[0] until none is left, and the last
    // workgroup to leave (ctr[1]) clears both counters for the next launch.  A task computes
    // exactly what the workgroup of that index computes in ctpwa_eval, so results do not
    // depend on which workgroup runs it; the job set can change without changing the launch.
    o << "extern \"C\" __global__ void " << lb << " ctpwa_eval_group_persist(const EvalJob* __restrict__ jobs,\n"
         "    const int* __restrict__ hdr, unsigned int* __restrict__ ctr)\n{\n"
      << decl.str() <<
         "  __shared__ int task;\n"
         "  const int njobs = hdr[0], total = hdr[1];\n"
         "  for (;;) {\n"
         "    __syncthreads();                  // the previous task's reads of red / lbuf are done\n"
         "    if (threadIdx.x == 0) task = (int)atomicAdd(&ctr[0], 1u);\n"
         "    __syncthreads();\n"
         "    const int t = task;\n"
         "    if (t >= total) break;\n"
         "    int lo = 0, hi = njobs - 1;\n"
         "    while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (jobs[mid].wg_begin <= t) lo = mid; else hi = mid - 1; }\n"
         "    const EvalJob J = jobs[lo];\n"
         "    eval_body<true>(J.tab, J.sl, J.inv, J.invf, J.wts, J.part, J.wg_cls, J.wg_rng, J.ld, J.nwg, J.w1, J.w2, J.pw1, J.pw2,\n"
         "              t - J.wg_begin, red, lbuf, J.cf);\n"
         "  }\n"
         "  if (threadIdx.x == 0) {\n"
         "    __threadfence();\n"
         "    if (atomicAdd(&ctr[1], 1u) == gridDim.x - 1) { atomicExch(&ctr[0], 0u); atomicExch(&ctr[1], 0u); }\n"
         "  }\n}\n";
    if (!M.customs.empty()) {
        // Custom factors and their derivatives for every stored event at the current
        // parameter point, rows [c * NCF + (r, i, dr[0..NCV), di[0..NCV))] of cfb.  A kernel
        // of its own: the complex-dual expression code then never shares registers with
        // the fused pass (which only loads the results).
        o << "extern \"C\" __global__ void __launch_bounds__(256) ctpwa_custom(\n"
             "    const double* __restrict__ tab, const double* __restrict__ inv, const float* __restrict__ invf,\n"
             "    long long ld, double* __restrict__ cfb)\n{\n"
             "  const double* tb = tab;\n"
             "  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < ld; e += (long long)gridDim.x * 256) {\n";
        for (size_t c = 0; c < M.customs.size(); ++c) {
            const CustomSpec& C = M.customs[c];
            size_t ti = 0;
            std::string code = custom_code(C.expr, ti, C, M.bvars[C.block]);
            if (ti != C.expr.size()) throw std::runtime_error("Custom expression has trailing tokens");
            const long long r0 = (long long)c * M.ncf();
            o << "    {\n      const double xm = " << inv_load(C.inv_m, "e") << ", xq = " << inv_load(C.inv_q, "e") << ";\n";
            for (int j = 0; j < 8; ++j)
                if (C.inv_d[j] >= 0) o << "      const double xd" << j << " = " << inv_load(C.inv_d[j], "e") << ";\n";
            o
              << "      const CD F = " << code << ";\n"
              << "      cfb[" << r0 << "LL * ld + e] = F.r; cfb[" << r0 + 1 << "LL * ld + e] = F.i;\n"
              << "      for (int k = 0; k < NCV; ++k) { cfb[(" << r0 + 2 << "LL + k) * ld + e] = F.dr[k]; cfb[(" << r0 + 2
              << "LL + NCV + k) * ld + e] = F.di[k]; }\n    }\n";
        }
        o << "  }\n}\n";
        // per-event amplitudes (without couplings) or intensities of a range of events:
        // models with Custom lineshapes evaluate them with the same generated core
        o << "extern \"C\" __global__ void __launch_bounds__(256) ctpwa_aux(\n"
             "    const double* __restrict__ tab, const void* __restrict__ slv, const double* __restrict__ inv,\n"
             "    const float* __restrict__ invf, long long ld, long long e0, long long n, double* __restrict__ out,\n"
             "    int intensity, const double* __restrict__ cfb)\n{\n";
        o << "  const " << SLT << "* __restrict__ sl = (const " << SLT << "*)slv;\n  const double* tb = tab;\n";
        emit_coefs("  ", false);
        o << "  const long long ie = (long long)blockIdx.x * 256 + threadIdx.x;\n  if (ie >= n) return;\n"
             "  const long long e = e0 + ie;\n  {\n";
        for (int u = 0; u < NU; ++u)
            for (int p = 0; p < NP; ++p) {
                std::string nm = std::to_string(u) + "_" + std::to_string(p);
                if (p >= NPS) {
                    o << mirror(u, p, "    ");
                    continue;
                }
                o << "    const " << SLT << " g" << nm << " = sl[" << (long long)u * NPS + p << "LL * ld + e];\n";
                o << "    const double2 s" << nm << " = " << sl_conv("g" + nm) << ";\n";
            }
        for (size_t i = 0; i < M.inv.size(); ++i)
            if (used[i]) o << "    const double x" << i << " = " << inv_load(i, "e") << ";\n";
        emit_core(o);
        o << "    if (intensity) {\n";
        for (int u = 0; u < NU; ++u) o << "      double y" << u << "r = 0.0, y" << u << "i = 0.0;\n";
        for (int a = 0; a < NA; ++a) {
            int b = M.amps[a].block, u = M.amps[a].u;
            o << "      y" << u << "r += c" << a << "r * br" << b << " - c" << a << "i * bi" << b << "; y" << u
              << "i += c" << a << "r * bi" << b << " + c" << a << "i * br" << b << ";\n";
        }
        o << "      double I = 0.0;\n";
        for (int p = 0; p < NP; ++p) {
            o << "      { double sr = 0.0, si = 0.0;\n";
            for (int u = 0; u < NU; ++u)
                o << "        sr += y" << u << "r * s" << u << "_" << p << ".x - y" << u << "i * s" << u << "_" << p
                  << ".y; si += y" << u << "r * s" << u << "_" << p << ".y + y" << u << "i * s" << u << "_" << p << ".x;\n";
            o << "        I += sr * sr + si * si; }\n";
        }
        o << "      out[ie] = I;\n    } else {\n";
        for (int a = 0; a < NA; ++a) {
            int b = M.amps[a].block, u = M.amps[a].u;
            o << "      { const double K = tb[OFFR + " << (long long)M.amps[a].coef_base << "LL * REC];\n";
            for (int p = 0; p < NP; ++p) {
                o << "        { const double fr = K * br" << b << ", fi = K * bi" << b << ";\n"
                  << "          const long long oo = ((ie * " << NP << " + " << p << ") * " << NA << " + " << a << ") * 2;\n"
                  << "          out[oo] = fr * s" << u << "_" << p << ".x - fi * s" << u << "_" << p << ".y;\n"
                  << "          out[oo + 1] = fr * s" << u << "_" << p << ".y + fi * s" << u << "_" << p << ".x; }\n";
            }
            o << "      }\n";
        }
        o << "    }\n  }\n}\n";
    }
    return o.str();
}

}  // namespace ctpwa
