// Source generator for the Hessian rows kernel (replaces the per-event part of
// analysis::getHessian, src/Analysis.cu:2689-3378, src/AmpGen.cu:3020-3540,
// src/ComputeHessian.cu:411-1170).
//
// Like the fused likelihood kernel, the code is specialised to one model: every
// denominator, block product and amplitude is a straight line of register arithmetic,
// so an event's work has no data-dependent loops and no dependent table loads.  Per
// event (one thread) it writes
//   J      the Jacobian rows: for every polarisation p the complex columns
//          A_a,p = F_a sl_u(a),p (amplitudes) and dS_p/dtheta_j (free parameters),
//          real parts in columns [0, C), imaginary parts in [C, 2C), column-major over
//          the chunk's rows p * n + r; FP64, or FP32 in the mixed form (the Gram then
//          runs on the FP32 matrix cores with FP64 accumulation across tiles);
//   U      u = dI/dx (x = Re v, Im v, theta) for data and background (its weighted Gram
//          is the -w/I^2 part of the second derivative of w ln I);
//   W      alpha = w / I and beta = -w / I^2 (data, background) or 1 and 0 (phase space);
// and sums over each wave's 64 events, in LDS, of the alpha-weighted second-part scalars
// <S, dA_a/dtheta_i>, sum_a Re(v_a d2F_a/dtheta_i dtheta_k conj(T_u)), alpha I, and of u.
//
// With psi_d = 1 / D_d, the log-derivatives L_i = sum_d psi_d dD_d/dtheta_i and
// M_ik = sum_d (psi_d d2D_d - psi_d^2 dD_i dD_k) give the block product's derivatives
// P1 = -dPsi = Psi L, P2 = d2Psi = Psi (L L^T - M); Custom factors multiply in by the
// product rule from their per-event value, first and second derivative rows.
#include <algorithm>
#include <sstream>
#include <stdexcept>
#include <string>

#include "model.hpp"

namespace ctpwa {

static std::string DH(double x) {
    std::ostringstream s;
    s.precision(17);
    s << x;
    std::string r = s.str();
    if (r.find_first_of(".eEn") == std::string::npos) r += ".0";
    return r;
}

// fixed-order column sums over a wave's 64 lanes: every lane hands in the same sequence
// of values; 16 at a time are staged in the wave's LDS slab (rows padded to 17 doubles),
// lane 4k + q adds column k of lanes 16q .. 16q + 15 in lane order and two exchanges
// join the quarters.  The wave's sum of column c goes to out[c * ld].
static const char* HR_PREAMBLE = R"(
__device__ __forceinline__ double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0); r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0); return __builtin_fma(r, e, r);
}
struct WaveColSums {
  double (*slab)[17];
  double* out;
  long long ld;
  int lane, n, col;
  __device__ __forceinline__ void flush() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int k = lane >> 2, q = lane & 3;
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < 16; ++t) s += slab[16 * q + t][k];
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    if (q == 0 && k < n) out[(long long)(col + k) * ld] = s;
    col += n;
    n = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __device__ __forceinline__ void put(double v) {
    slab[lane][n] = v;
    if (++n == 16) flush();
  }
};
)";

// Blocks are visited in an order that keeps the blocks of one parameter set together
// (the per-event derivative columns of a parameter then live in registers only from its
// first to its last block); the summed columns are handed to the wave sums in that
// order, and perm (when given) receives, for every summed column in put order, its
// index in the assembly's order [Z (plan block order), alpha I, u].
std::string generate_hess_kernel(const Model_& M, std::vector<int>* perm) {
    const Plan& P = M.P;
    const int NP = P.npol, NA = P.namp, NT = P.ntheta;
    const int C = NA + NT;
    const int NUX = 2 * NA + NT;
    std::ostringstream o;
    o << "// generated: Hessian rows kernel\n";
    o << "#define OFFR " << M.off_rec() << "LL\n#define OFFV " << M.off_v() << "LL\n#define REC " << Model_::REC << "\n";
    o << HR_PREAMBLE;
    const char* SLT = M.mixed ? "float2" : "double2";
    const char* JT = M.mixed ? "float" : "double";
    const std::vector<int> imap = M.imap();
    // component of the second derivative (i, k) in a coefficient record
    auto hidx = [](int i, int k) { return 1 + Model_::NLOC + i * Model_::NLOC + k; };
    o << "extern \"C\" __global__ void __launch_bounds__(256) ctpwa_hess_rows(\n"
         "    const double* __restrict__ tab, const void* __restrict__ slv, const double* __restrict__ inv,\n"
         "    const float* __restrict__ invf, long long ld, const double* __restrict__ wts, double wconst, int cls,\n"
         "    long long e0, long long n, " << JT << "* __restrict__ J, long long ldJ, long long ldc,\n"
         "    double* __restrict__ U, double* __restrict__ W, double* __restrict__ cs, long long cs_ld,\n"
         "    const double* __restrict__ cfb2, long long cld)\n{\n";
    o << "  __shared__ double slab[4][64][17];\n"
         "  const long long r0 = (long long)blockIdx.x * 256 + threadIdx.x;\n"
         "  const bool live = r0 < n;\n"
         // a thread past the chunk's end takes its last event and stores the same values as
         // that event's own thread: its stores need no guard.  Guarded stores (a divergent
         // branch each) split a large model's straight-line code into hundreds of blocks,
         // around which the register allocator spilled: 17 % of the 149-amplitude Hessian
         "  const long long r = live ? r0 : n - 1;\n"
         "  const int wv = threadIdx.x >> 6;\n"
         "  WaveColSums cs_{slab[wv], cs + (long long)blockIdx.x * 4 + wv, cs_ld, (int)(threadIdx.x & 63), 0, 0};\n"
         "  const " << SLT << "* __restrict__ sl = (const " << SLT << "*)slv;\n";
    // Every per-event input is loaded once, at the start, before any store: on this
    // target a wait for a load issued after stores also waits for those stores, so loads
    // placed among the Jacobian stores would serialise the kernel on store completion.
    // Each phase reads the wave-uniform coefficients at its own offset of the table
    // pointer, a zero passed through an empty asm statement: the compiler then reloads
    // them (scalar loads) where the second phase needs them instead of keeping the first
    // phase's values live in registers.
    auto phase_vars = [&](int ph) {
        o << "  long long z_" << ph << " = 0;\n  asm volatile(\"\" : \"+s\"(z_" << ph << "));\n"
          << "  const double* __restrict__ tb_" << ph << " = tab + z_" << ph << ";\n";
    };
    std::vector<char> xused(M.inv.size(), 0);
    for (const DenSpec& Dn : M.dens) {
        for (const Term& t : Dn.re) if (t.inv >= 0) xused[t.inv] = 1;
        for (const Term& t : Dn.im) if (t.inv >= 0) xused[t.inv] = 1;
    }
    o << "  const long long e = e0 + r;\n"
         "  const double wload = (cls != 0 && wts) ? wts[e] : wconst;\n";
    for (size_t i = 0; i < M.inv.size(); ++i) {
        if (!xused[i]) continue;
        if (imap[i] >= 0) o << "  const double x" << i << " = inv[" << imap[i] << "LL * ld + e];\n";
        else o << "  const double x" << i << " = (double)invf[" << (-imap[i] - 1) << "LL * ld + e];\n";
    }
    // stored polarisations per SL tensor; a polarisation past them is the signed
    // conjugate of its mirror image
    const int NPS = M.npstore();
    for (size_t u = 0; u < M.us.size(); ++u)
        for (int p = 0; p < NP; ++p) {
            std::string nm = std::to_string(u) + "_" + std::to_string(p);
            if (p >= NPS) {
                const std::string mm = std::to_string(u) + "_" + std::to_string(NP - 1 - p);
                const bool neg = M.psign[u] < 0;
                o << "  const double s" << nm << "x = " << (neg ? "-" : "") << "s" << mm << "x, s" << nm << "y = " << (neg ? "" : "-")
                  << "s" << mm << "y;\n";
                continue;
            }
            o << "  const " << SLT << " g" << nm << " = sl[" << (long long)u * NPS + p << "LL * ld + e];\n";
            o << "  const double s" << nm << "x = (double)g" << nm << ".x, s" << nm << "y = (double)g" << nm << ".y;\n";
        }
    // Custom factor rows (value, first and second derivatives) of the event
    for (size_t c = 0; c < M.customs.size(); ++c)
        for (int comp = 0; comp < (M.customs.empty() ? 0 : M.ncf2()); ++comp)
            o << "  const double cf" << c << "_" << comp << " = cfb2[" << (long long)c * M.ncf2() + comp << "LL * cld + e];\n";
    auto xv = [&](int, int i) { return "x" + std::to_string(i); };
    auto rec = [&](int ph, long long rr, int comp) {
        std::ostringstream q;
        q << "tb_" << ph << "[OFFR + " << rr << "LL * REC + " << comp << "]";
        return q.str();
    };
    const int ncf2 = M.customs.empty() ? 0 : M.ncf2();
    const int ncv = M.ncv();
    const int nh = ncv * (ncv + 1) / 2;
    auto cf = [&](int, int c, int comp) { return "cf" + std::to_string(c) + "_" + std::to_string(comp); };
    (void)ncf2;
    // the propagators of block b (psi_d = pr + i pi) and the block product (br + i bi)
    auto emit_block_value = [&](int ph, int b, const char* ind) {
        const std::vector<int>& ds = M.block_dens[b];
        for (int d : ds) {
            const DenSpec& Dn = M.dens[d];
            auto term = [&](const Term& t) {
                std::string c = (t.coef < 0) ? DH(t.cval) : rec(ph, Dn.coef_base + t.coef, 0);
                if (t.inv < 0) return c;
                return c + " * " + xv(ph, t.inv);
            };
            o << ind << "double dr" << d << " = 0.0, di" << d << " = 0.0;\n";
            for (const Term& t : Dn.re) o << ind << "dr" << d << " += " << term(t) << ";\n";
            for (const Term& t : Dn.im) o << ind << "di" << d << " += " << term(t) << ";\n";
            if (Dn.stat) {
                o << ind << "const double pr" << d << " = dr" << d << ", pi" << d << " = di" << d << ";\n";
            } else {
                o << ind << "const double q" << d << " = frcp(dr" << d << " * dr" << d << " + di" << d << " * di" << d << ");\n";
                o << ind << "const double pr" << d << " = dr" << d << " * q" << d << ", pi" << d << " = di" << d << " * q" << d << ";\n";
            }
        }
        o << ind << "double br = 1.0, bi = 0.0;\n";
        for (int d : ds)
            o << ind << "{ const double tr = br * pr" << d << " - bi * pi" << d << "; bi = br * pi" << d << " + bi * pr" << d
              << "; br = tr; }\n";
    };
    auto sload = [&](int, int, int, const char*) {};

    // ---- phase 1: block products, amplitude columns, S, I, weights -------------------
    phase_vars(1);
    for (int p = 0; p < NP; ++p) o << "  double S" << p << "r = 0.0, S" << p << "i = 0.0;\n";
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const PBlock& B = P.blocks[b];
        if (B.nsl == 0) continue;
        o << "  {\n";
        emit_block_value(1, (int)b, "    ");
        for (int c : M.block_customs[b])
            o << "    { const double fr = " << cf(1, c, 0) << ", fi = " << cf(1, c, 1) << ";\n"
              << "      const double tr = br * fr - bi * fi; bi = br * fi + bi * fr; br = tr; }\n";
        std::vector<char> loaded(M.us.size(), 0);
        for (int cc = 0; cc < B.nsl; ++cc) {
            const int a = B.site + cc, u = M.amps[a].u;
            if (!loaded[u])
                for (int p = 0; p < NP; ++p) sload(1, u, p, "    ");
            loaded[u] = 1;
            o << "    { const double K = " << rec(1, M.amps[a].coef_base, 0) << ";\n"
              << "      const double Fr = K * br, Fi = K * bi;\n"
              << "      const double vr = tb_1[OFFV + " << 2 * a << "], vi = tb_1[OFFV + " << 2 * a + 1 << "];\n";
            for (int p = 0; p < NP; ++p) {
                std::string sn = "s" + std::to_string(u) + "_" + std::to_string(p);
                o << "      { const double ar = Fr * " << sn << "x - Fi * " << sn << "y, ai = Fr * " << sn << "y + Fi * " << sn << "x;\n"
                  << "        S" << p << "r += vr * ar - vi * ai; S" << p << "i += vr * ai + vi * ar;\n"
                  << "        J[" << a << "LL * ldJ + " << p << "LL * n + r] = (" << JT << ")ar; J[" << C + a << "LL * ldJ + " << p
                  << "LL * n + r] = (" << JT << ")ai; }\n";
            }
            o << "    }\n";
        }
        o << "  }\n";
    }
    o << "  double I = 0.0;\n";
    for (int p = 0; p < NP; ++p) o << "  I += S" << p << "r * S" << p << "r + S" << p << "i * S" << p << "i;\n";
    o << "  const double we = (cls == 0) ? 1.0 : wload;\n"
         "  double alpha, beta;\n"
         "  if (cls == 0) { alpha = 1.0; beta = 0.0; }\n"
         "  else if (I == 0.0) { alpha = 0.0; beta = 0.0; }\n"
         "  else { const double iI = frcp(I); alpha = we * iI; beta = -alpha * iI; }\n"
         "  W[r] = alpha; W[ldc + r] = beta;\n"
         "  const bool ust = cls != 0;   // wave-uniform: a scalar branch around each u store\n";

    // ---- phase 2: u, derivative columns and second-part scalars --------------------
    // canonical column of every summed value: Z in plan block order, then alpha I, then u
    std::vector<int> zoff(P.blocks.size(), 0);
    int NZ = 0;
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        zoff[b] = NZ;
        if (P.blocks[b].nsl == 0) continue;
        const int nl = (int)M.bvars[b].size();
        NZ += P.blocks[b].nsl * 2 * nl + nl * (nl + 1) / 2;
    }
    const int col_I = NZ, col_u0 = NZ + 1;
    std::vector<int> order_cols;
    auto put = [&](const std::string& v, int col, const char* ind) {
        o << ind << "cs_.put(live ? " << v << " : 0.0);\n";
        order_cols.push_back(col);
    };
    auto ustore = [&](const std::string& v, int x, const char* ind) {
        put(v, col_u0 + x, ind);
        o << ind << "if (ust) U[" << x << "LL * ldc + r] = " << v << ";\n";
    };
    phase_vars(2);
    put("alpha * I", col_I, "  ");
    // block order: blocks with waves, grouped by their parameter sets
    std::vector<int> border;
    for (size_t b = 0; b < P.blocks.size(); ++b) if (P.blocks[b].nsl > 0) border.push_back((int)b);
    std::stable_sort(border.begin(), border.end(), [&](int x, int y) {
        std::vector<int> vx = M.bvars[x], vy = M.bvars[y];
        std::sort(vx.begin(), vx.end());
        std::sort(vy.begin(), vy.end());
        return vx < vy;
    });
    std::vector<int> last(NT, -1);
    for (size_t k = 0; k < border.size(); ++k)
        for (int j : M.bvars[border[k]]) last[j] = (int)k;
    std::vector<char> open_(NT, 0);
    for (size_t kb = 0; kb < border.size(); ++kb) {
        const int b = border[kb];
        const PBlock& B = P.blocks[b];
        const std::vector<int>& vars = M.bvars[b];
        const int nl = (int)vars.size();
        for (int j : vars)
            if (!open_[j]) {
                open_[j] = 1;
                o << "  double ut" << j << " = 0.0;\n";
                for (int p = 0; p < NP; ++p) o << "  double dS" << j << "_" << p << "r = 0.0, dS" << j << "_" << p << "i = 0.0;\n";
            }
        o << "  {\n";
        emit_block_value(2, b, "    ");
        o << "    double psr = br, psi_ = bi;\n";
        if (nl > 0) {
            // log-derivatives of the propagators: l_i = psi dD_i, M_ik = sum (psi d2D_ik - l_i l_k)
            for (int i = 0; i < nl; ++i) o << "    double Lr" << i << " = 0.0, Li" << i << " = 0.0;\n";
            for (int i = 0; i < nl; ++i)
                for (int k = i; k < nl; ++k) o << "    double Mr" << i << "_" << k << " = 0.0, Mi" << i << "_" << k << " = 0.0;\n";
            for (int d : M.block_dens[b]) {
                const DenSpec& Dn = M.dens[d];
                if (Dn.stat) continue;
                bool any = false;
                for (const Term& t : Dn.re) if (t.coef >= 0) any = true;
                for (const Term& t : Dn.im) if (t.coef >= 0) any = true;
                if (!any) continue;
                auto dsum = [&](const char* acc, int comp, const char* ind) {
                    for (const Term& t : Dn.re)
                        if (t.coef >= 0)
                            o << ind << acc << "r += " << rec(2, Dn.coef_base + t.coef, comp) << (t.inv >= 0 ? " * " + xv(2, t.inv) : "") << ";\n";
                    for (const Term& t : Dn.im)
                        if (t.coef >= 0)
                            o << ind << acc << "i += " << rec(2, Dn.coef_base + t.coef, comp) << (t.inv >= 0 ? " * " + xv(2, t.inv) : "") << ";\n";
                };
                o << "    {\n";
                // dD_i = gr - i gi; l_i = (pr + i pi)(gr - i gi)
                for (int i = 0; i < nl; ++i) {
                    o << "      double g" << i << "r = 0.0, g" << i << "i = 0.0;\n";
                    dsum(("g" + std::to_string(i)).c_str(), 1 + i, "      ");
                    o << "      const double lr" << i << " = pr" << d << " * g" << i << "r + pi" << d << " * g" << i << "i, li" << i << " = pi" << d
                      << " * g" << i << "r - pr" << d << " * g" << i << "i;\n";
                    o << "      Lr" << i << " += lr" << i << "; Li" << i << " += li" << i << ";\n";
                }
                for (int i = 0; i < nl; ++i)
                    for (int k = i; k < nl; ++k) {
                        o << "      { double hr = 0.0, hi = 0.0;\n";
                        for (const Term& t : Dn.re)
                            if (t.coef >= 0)
                                o << "        hr += " << rec(2, Dn.coef_base + t.coef, hidx(i, k)) << (t.inv >= 0 ? " * " + xv(2, t.inv) : "") << ";\n";
                        for (const Term& t : Dn.im)
                            if (t.coef >= 0)
                                o << "        hi += " << rec(2, Dn.coef_base + t.coef, hidx(i, k)) << (t.inv >= 0 ? " * " + xv(2, t.inv) : "") << ";\n";
                        o << "        Mr" << i << "_" << k << " += pr" << d << " * hr + pi" << d << " * hi - (lr" << i << " * lr" << k << " - li" << i
                          << " * li" << k << ");\n"
                          << "        Mi" << i << "_" << k << " += pi" << d << " * hr - pr" << d << " * hi - (lr" << i << " * li" << k << " + li" << i
                          << " * lr" << k << "); }\n";
                    }
                o << "    }\n";
            }
            // P1 = Psi L, P2 = Psi (L L - M)
            for (int i = 0; i < nl; ++i)
                for (int k = i; k < nl; ++k)
                    o << "    double P2r" << i << "_" << k << ", P2i" << i << "_" << k << ";\n"
                      << "    { const double ar = Lr" << i << " * Lr" << k << " - Li" << i << " * Li" << k << " - Mr" << i << "_" << k << ", ai = Lr" << i
                      << " * Li" << k << " + Li" << i << " * Lr" << k << " - Mi" << i << "_" << k << ";\n"
                      << "      P2r" << i << "_" << k << " = psr * ar - psi_ * ai; P2i" << i << "_" << k << " = psr * ai + psi_ * ar; }\n";
            for (int i = 0; i < nl; ++i)
                o << "    double P1r" << i << " = psr * Lr" << i << " - psi_ * Li" << i << ", P1i" << i << " = psr * Li" << i << " + psi_ * Lr" << i << ";\n";
        }
        // Custom factors: d2(Psi F) = P2 F - P1_i F_k - P1_k F_i + Psi F_ik; -(Psi F)_i = P1_i F - Psi F_i
        for (int c : M.block_customs[b]) {
            o << "    {\n      const double fr = " << cf(2, c, 0) << ", fi = " << cf(2, c, 1) << ";\n";
            for (int i = 0; i < nl; ++i)
                o << "      const double fr" << i << " = " << cf(2, c, 2 + i) << ", fi" << i << " = " << cf(2, c, 2 + ncv + i) << ";\n";
            for (int i = 0; i < nl; ++i)
                for (int k = i; k < nl; ++k) {
                    const int h = i * ncv - (i * (i - 1)) / 2 + (k - i);
                    o << "      { const double hr = " << cf(2, c, 2 + 2 * ncv + h) << ", hi = " << cf(2, c, 2 + 2 * ncv + nh + h) << ";\n"
                      << "        const double nr = P2r" << i << "_" << k << " * fr - P2i" << i << "_" << k << " * fi - (P1r" << i << " * fr" << k
                      << " - P1i" << i << " * fi" << k << ") - (P1r" << k << " * fr" << i << " - P1i" << k << " * fi" << i << ") + (psr * hr - psi_ * hi);\n"
                      << "        const double ni = P2r" << i << "_" << k << " * fi + P2i" << i << "_" << k << " * fr - (P1r" << i << " * fi" << k
                      << " + P1i" << i << " * fr" << k << ") - (P1r" << k << " * fi" << i << " + P1i" << k << " * fr" << i << ") + (psr * hi + psi_ * hr);\n"
                      << "        P2r" << i << "_" << k << " = nr; P2i" << i << "_" << k << " = ni; }\n";
                }
            for (int i = 0; i < nl; ++i)
                o << "      { const double nr = P1r" << i << " * fr - P1i" << i << " * fi - (psr * fr" << i << " - psi_ * fi" << i << ");\n"
                  << "        const double ni = P1r" << i << " * fi + P1i" << i << " * fr - (psr * fi" << i << " + psi_ * fr" << i << ");\n"
                  << "        P1r" << i << " = nr; P1i" << i << " = ni; }\n";
            o << "      { const double tr = psr * fr - psi_ * fi; psi_ = psr * fi + psi_ * fr; psr = tr; }\n    }\n";
        }
        for (int i = 0; i < nl; ++i)
            for (int k = i; k < nl; ++k) o << "    double R" << i << "_" << k << " = 0.0;\n";
        std::vector<char> loaded(M.us.size(), 0);
        for (int cc = 0; cc < B.nsl; ++cc) {
            const int a = B.site + cc, u = M.amps[a].u;
            const long long ra = M.amps[a].coef_base;
            if (!loaded[u]) {
                for (int p = 0; p < NP; ++p) sload(2, u, p, "    ");
                // T_u = sum_p conj(sl_u,p) S_p
                o << "    double t" << u << "r = 0.0, t" << u << "i = 0.0;\n";
                for (int p = 0; p < NP; ++p) {
                    std::string sn = "s" + std::to_string(u) + "_" + std::to_string(p);
                    o << "    t" << u << "r += " << sn << "x * S" << p << "r + " << sn << "y * S" << p << "i; t" << u << "i += " << sn << "x * S" << p
                      << "i - " << sn << "y * S" << p << "r;\n";
                }
            }
            loaded[u] = 1;
            o << "    {\n      const double K = " << rec(2, ra, 0) << ";\n"
              << "      const double vr = tb_2[OFFV + " << 2 * a << "], vi = tb_2[OFFV + " << 2 * a + 1 << "];\n"
              << "      const double Fr = K * psr, Fi = K * psi_;\n"
              // g_a = conj(F_a) T_u: dI/dRe v_a = 2 Re g, dI/dIm v_a = 2 Im g
              << "      const double gr = Fr * t" << u << "r + Fi * t" << u << "i, gi = Fr * t" << u << "i - Fi * t" << u << "r;\n";
            ustore("2.0 * gr", a, "      ");
            ustore("2.0 * gi", NA + a, "      ");
            for (int i = 0; i < nl; ++i) {
                // dF_i = K_i Psi - K P1_i; q = dF_i conj(T_u) = <S, dA_a/dtheta_i>
                o << "      const double Kg" << i << " = " << rec(2, ra, 1 + i) << ";\n"
                  << "      const double dFr" << i << " = Kg" << i << " * psr - K * P1r" << i << ", dFi" << i << " = Kg" << i << " * psi_ - K * P1i" << i << ";\n"
                  << "      const double qr" << i << " = dFr" << i << " * t" << u << "r + dFi" << i << " * t" << u << "i, qi" << i << " = dFi" << i << " * t" << u
                  << "r - dFr" << i << " * t" << u << "i;\n";
                put("alpha * qr" + std::to_string(i), zoff[b] + cc * 2 * nl + 2 * i, "      ");
                put("alpha * qi" + std::to_string(i), zoff[b] + cc * 2 * nl + 2 * i + 1, "      ");
                o << "      ut" << vars[i] << " += 2.0 * (vr * qr" << i << " - vi * qi" << i << ");\n";
                o << "      { const double vdr = vr * dFr" << i << " - vi * dFi" << i << ", vdi = vr * dFi" << i << " + vi * dFr" << i << ";\n";
                for (int p = 0; p < NP; ++p) {
                    std::string sn = "s" + std::to_string(u) + "_" + std::to_string(p);
                    o << "        dS" << vars[i] << "_" << p << "r += vdr * " << sn << "x - vdi * " << sn << "y; dS" << vars[i] << "_" << p << "i += vdr * " << sn
                      << "y + vdi * " << sn << "x;\n";
                }
                o << "      }\n";
            }
            // R_ik += Re(v d2F_ik conj(T_u)), d2F_ik = K_ik Psi - K_i P1_k - K_k P1_i + K P2_ik
            for (int i = 0; i < nl; ++i)
                for (int k = i; k < nl; ++k)
                    o << "      { const double Kh = " << rec(2, ra, hidx(i, k)) << ";\n"
                      << "        const double fr = Kh * psr - Kg" << i << " * P1r" << k << " - Kg" << k << " * P1r" << i << " + K * P2r" << i << "_" << k << ";\n"
                      << "        const double fi = Kh * psi_ - Kg" << i << " * P1i" << k << " - Kg" << k << " * P1i" << i << " + K * P2i" << i << "_" << k << ";\n"
                      << "        R" << i << "_" << k << " += (vr * fr - vi * fi) * t" << u << "r + (vr * fi + vi * fr) * t" << u << "i; }\n";
            o << "    }\n";
        }
        {
            int z = zoff[b] + B.nsl * 2 * nl;
            for (int i = 0; i < nl; ++i)
                for (int k = i; k < nl; ++k) put("alpha * R" + std::to_string(i) + "_" + std::to_string(k), z++, "    ");
        }
        o << "  }\n";
        // derivative columns of the parameters whose last block this was
        for (int j : vars)
            if (last[j] == (int)kb) {
                for (int p = 0; p < NP; ++p)
                    o << "  J[" << NA + j << "LL * ldJ + " << p << "LL * n + r] = (" << JT << ")dS" << j << "_" << p << "r; J[" << C + NA + j
                      << "LL * ldJ + " << p << "LL * n + r] = (" << JT << ")dS" << j << "_" << p << "i;\n";
                ustore("ut" + std::to_string(j), 2 * NA + j, "  ");
            }
    }
    // parameters of no block with waves: zero columns
    for (int j = 0; j < NT; ++j)
        if (last[j] < 0) {
            for (int p = 0; p < NP; ++p)
                o << "  J[" << NA + j << "LL * ldJ + " << p << "LL * n + r] = (" << JT << ")0.0; J[" << C + NA + j << "LL * ldJ + "
                  << p << "LL * n + r] = (" << JT << ")0.0;\n";
            ustore("0.0", 2 * NA + j, "  ");
        }
    o << "  if (cs_.n > 0) cs_.flush();\n}\n";
    if ((int)order_cols.size() != NZ + 1 + NUX) throw std::runtime_error("Hessian generator: summed-column count mismatch");
    if (perm) *perm = order_cols;
    return o.str();
}

}  // namespace ctpwa
