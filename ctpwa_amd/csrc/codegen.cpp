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
//   accT_j += -2 Re sum_{b,d} Z_b Psi_b Psi_d dD_d/dtheta_j,  Z_b = sum_{a in b} c_a conj(T_u(a))
//   accV   += I (phase space) | w_e ln I (data, background)
// then a deterministic wave/LDS reduction to one partial row per workgroup.
#include <sstream>
#include <string>

#include "model.hpp"

namespace ctpwa {

static std::string D(double x) {
    std::ostringstream s;
    s.precision(17);
    s << x;
    std::string r = s.str();
    if (r.find_first_of(".eEn") == std::string::npos) r += ".0";
    return r;
}

// local variables of block b that a denominator coefficient can depend on
static bool coef_depends(const Model_& M, const DenSpec& d, int coef, int local) {
    const PBlock& B = M.P.blocks[d.block];
    const PRes& R = B.res[d.res];
    int slot = M.bvars[d.block][local];
    bool own_mass = !R.slots.empty() && R.slots[0] == slot;
    if ((R.model == M_BWR || R.model == M_BW || R.model == M_FLATTE) && coef == 0) return own_mass;
    return true;
}

std::string generate_eval_kernel(const Model_& M) {
    const Plan& P = M.P;
    int NU = (int)M.us.size(), NP = P.npol, NA = P.namp, NT = P.ntheta;
    int NACC = 1 + 2 * NA + NT;
    std::ostringstream o;
    o << "// generated: fused likelihood kernel\n";
    o << "#define OFFC " << M.off_c() << "LL\n#define OFFR " << M.off_rec() << "LL\n#define REC " << Model_::REC << "\n";
    o << "extern \"C\" __global__ void __launch_bounds__(256) ctpwa_eval(\n"
         "    const double* __restrict__ tab, const double2* __restrict__ sl, const double* __restrict__ inv,\n"
         "    const double* __restrict__ wts, double* __restrict__ part, const int* __restrict__ wg_cls,\n"
         "    const long long* __restrict__ wg_rng, long long ld, int nwg, double w1, double w2, int pw1, int pw2)\n{\n";
    o << "  __shared__ double red[4][" << NACC << "];\n";
    o << "  const int wg = blockIdx.x;\n  const int cls = wg_cls[wg];\n"
         "  const long long e0 = wg_rng[2 * wg], e1 = wg_rng[2 * wg + 1];\n"
         "  const double wc = (cls == 1) ? w1 : w2;\n  const int pw = (cls == 1) ? pw1 : ((cls == 2) ? pw2 : 0);\n"
         "  const double LOG_FLOOR = log(1e-10);\n";
    // coefficients
    for (int a = 0; a < NA; ++a)
        o << "  const double c" << a << "r = tab[OFFC + " << 2 * a << "], c" << a << "i = tab[OFFC + " << 2 * a + 1 << "];\n";
    for (size_t d = 0; d < M.dens.size(); ++d) {
        const DenSpec& Dn = M.dens[d];
        int nloc = (int)M.bvars[Dn.block].size();
        for (int c = 0; c < Dn.ncoef; ++c) {
            long long rec = Dn.coef_base + c;
            o << "  const double k" << d << "_" << c << " = tab[OFFR + " << rec << "LL * REC];\n";
            for (int i = 0; i < nloc; ++i)
                if (coef_depends(M, Dn, c, i))
                    o << "  const double g" << d << "_" << c << "_" << i << " = tab[OFFR + " << rec << "LL * REC + " << 1 + i << "];\n";
        }
    }
    o << "  double av = 0.0;\n";
    for (int a = 0; a < NA; ++a) o << "  double ax" << a << "r = 0.0, ax" << a << "i = 0.0;\n";
    for (int j = 0; j < NT; ++j) o << "  double at" << j << " = 0.0;\n";
    // which invariants are used
    std::vector<int> used(M.inv.size(), 0);
    for (const DenSpec& Dn : M.dens) {
        for (const Term& t : Dn.re) if (t.inv >= 0) used[t.inv] = 1;
        for (const Term& t : Dn.im) if (t.inv >= 0) used[t.inv] = 1;
    }
    o << "  for (long long e = e0 + threadIdx.x; e < e1; e += 256) {\n";
    for (int u = 0; u < NU; ++u)
        for (int p = 0; p < NP; ++p)
            o << "    const double2 s" << u << "_" << p << " = sl[" << (long long)u * NP + p << "LL * ld + e];\n";
    for (size_t i = 0; i < M.inv.size(); ++i)
        if (used[i]) o << "    const double x" << i << " = inv[" << i << "LL * ld + e];\n";
    auto term_expr = [&](int d, const Term& t) {
        std::string coef = (t.coef < 0) ? D(t.cval) : ("k" + std::to_string(d) + "_" + std::to_string(t.coef));
        if (t.inv < 0) return coef;
        return coef + " * x" + std::to_string(t.inv);
    };
    for (size_t d = 0; d < M.dens.size(); ++d) {
        const DenSpec& Dn = M.dens[d];
        o << "    const double dr" << d << " = ";
        for (size_t t = 0; t < Dn.re.size(); ++t) o << (t ? " + " : "") << term_expr((int)d, Dn.re[t]);
        if (Dn.re.empty()) o << "0.0";
        o << ";\n    const double di" << d << " = ";
        for (size_t t = 0; t < Dn.im.size(); ++t) o << (t ? " + " : "") << term_expr((int)d, Dn.im[t]);
        if (Dn.im.empty()) o << "0.0";
        o << ";\n    const double q" << d << " = 1.0 / (dr" << d << " * dr" << d << " + di" << d << " * di" << d << ");\n";
        o << "    const double pr" << d << " = dr" << d << " * q" << d << ", pi" << d << " = di" << d << " * q" << d << ";\n";
    }
    // block propagator products
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const std::vector<int>& ds = M.block_dens[b];
        if (P.blocks[b].nsl == 0) continue;
        if (ds.empty()) { o << "    const double br" << b << " = 1.0, bi" << b << " = 0.0;\n"; continue; }
        o << "    double br" << b << " = pr" << ds[0] << ", bi" << b << " = pi" << ds[0] << ";\n";
        for (size_t k = 1; k < ds.size(); ++k) {
            int d = ds[k];
            o << "    { const double tr = br" << b << " * pr" << d << " - bi" << b << " * pi" << d
              << "; bi" << b << " = br" << b << " * pi" << d << " + bi" << b << " * pr" << d << "; br" << b << " = tr; }\n";
        }
    }
    // Y_u = sum_a c_a Psi_b
    for (int u = 0; u < NU; ++u) o << "    double y" << u << "r = 0.0, y" << u << "i = 0.0;\n";
    for (int a = 0; a < NA; ++a) {
        int b = M.amps[a].block, u = M.amps[a].u;
        o << "    y" << u << "r += c" << a << "r * br" << b << " - c" << a << "i * bi" << b << "; y" << u << "i += c" << a
          << "r * bi" << b << " + c" << a << "i * br" << b << ";\n";
    }
    for (int p = 0; p < NP; ++p) {
        o << "    double S" << p << "r = 0.0, S" << p << "i = 0.0;\n";
        for (int u = 0; u < NU; ++u)
            o << "    S" << p << "r += y" << u << "r * s" << u << "_" << p << ".x - y" << u << "i * s" << u << "_" << p
              << ".y; S" << p << "i += y" << u << "r * s" << u << "_" << p << ".y + y" << u << "i * s" << u << "_" << p << ".x;\n";
    }
    o << "    const double I = ";
    for (int p = 0; p < NP; ++p) o << (p ? " + " : "") << "S" << p << "r * S" << p << "r + S" << p << "i * S" << p << "i";
    o << ";\n";
    for (int p = 0; p < NP; ++p) o << "    double w" << p << "r, w" << p << "i;\n";
    o << "    if (cls == 0) {\n      av += I;\n";
    for (int p = 0; p < NP; ++p) o << "      w" << p << "r = S" << p << "r; w" << p << "i = S" << p << "i;\n";
    o << "    } else {\n      const double we = pw ? wts[e] : wc;\n"
         "      if (I == 0.0) {\n        av += LOG_FLOOR * we;\n";
    for (int p = 0; p < NP; ++p) o << "        w" << p << "r = 0.0; w" << p << "i = 0.0;\n";
    o << "      } else {\n        av += log(I) * we;\n        const double f = we / I;\n";
    for (int p = 0; p < NP; ++p)
        o << "        w" << p << "r = fmin(fmax(S" << p << "r * f, -1e10), 1e10); w" << p << "i = fmin(fmax(S" << p
          << "i * f, -1e10), 1e10);\n";
    o << "      }\n    }\n";
    // T_u = sum_p conj(s_up) w_p
    for (int u = 0; u < NU; ++u) {
        o << "    double t" << u << "r = 0.0, t" << u << "i = 0.0;\n";
        for (int p = 0; p < NP; ++p)
            o << "    t" << u << "r += s" << u << "_" << p << ".x * w" << p << "r + s" << u << "_" << p << ".y * w" << p
              << "i; t" << u << "i += s" << u << "_" << p << ".x * w" << p << "i - s" << u << "_" << p << ".y * w" << p << "r;\n";
    }
    // accX_a += conj(Psi_b) T_u
    for (int a = 0; a < NA; ++a) {
        int b = M.amps[a].block, u = M.amps[a].u;
        o << "    ax" << a << "r += br" << b << " * t" << u << "r + bi" << b << " * t" << u << "i; ax" << a << "i += br" << b
          << " * t" << u << "i - bi" << b << " * t" << u << "r;\n";
    }
    // theta: Z_b = sum_{a in b} c_a conj(T_u);  R_d = Z_b Psi_b Psi_d
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const PBlock& B = P.blocks[b];
        if (B.nsl == 0 || M.block_dens[b].empty() || M.bvars[b].empty()) continue;
        o << "    {\n      double zr = 0.0, zi = 0.0;\n";
        for (int c = 0; c < B.nsl; ++c) {
            int a = B.site + c, u = M.amps[a].u;
            o << "      zr += c" << a << "r * t" << u << "r + c" << a << "i * t" << u << "i; zi += c" << a << "i * t" << u
              << "r - c" << a << "r * t" << u << "i;\n";
        }
        o << "      const double zbr = zr * br" << b << " - zi * bi" << b << ", zbi = zr * bi" << b << " + zi * br" << b << ";\n";
        for (int d : M.block_dens[b]) {
            const DenSpec& Dn = M.dens[d];
            o << "      {\n        const double rr = zbr * pr" << d << " - zbi * pi" << d << ", ri = zbr * pi" << d
              << " + zbi * pr" << d << ";\n";
            int nloc = (int)M.bvars[b].size();
            for (int i = 0; i < nloc; ++i) {
                std::ostringstream acc;
                bool any = false;
                for (const Term& t : Dn.re) {
                    if (t.coef < 0 || !coef_depends(M, Dn, t.coef, i)) continue;
                    acc << (any ? " + " : "") << "g" << d << "_" << t.coef << "_" << i << (t.inv >= 0 ? " * x" + std::to_string(t.inv) : "") << " * rr";
                    any = true;
                }
                for (const Term& t : Dn.im) {
                    if (t.coef < 0 || !coef_depends(M, Dn, t.coef, i)) continue;
                    acc << (any ? " + " : "") << "g" << d << "_" << t.coef << "_" << i << (t.inv >= 0 ? " * x" + std::to_string(t.inv) : "") << " * ri";
                    any = true;
                }
                if (any) o << "        at" << M.bvars[b][i] << " -= 2.0 * (" << acc.str() << ");\n";
            }
            o << "      }\n";
        }
        o << "    }\n";
    }
    o << "  }\n";
    // reduction
    o << "  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n";
    auto red = [&](const std::string& v, int k) {
        o << "  { double s = " << v << "; for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);"
             " if (lane == 0) red[wave][" << k << "] = s; }\n";
    };
    red("av", 0);
    for (int a = 0; a < NA; ++a) { red("ax" + std::to_string(a) + "r", 1 + 2 * a); red("ax" + std::to_string(a) + "i", 2 + 2 * a); }
    for (int j = 0; j < NT; ++j) red("at" + std::to_string(j), 1 + 2 * NA + j);
    o << "  __syncthreads();\n";
    o << "  for (int k = threadIdx.x; k < " << NACC << "; k += 256)\n"
         "    part[(long long)k * nwg + wg] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);\n";
    o << "}\n";
    return o.str();
}

}  // namespace ctpwa
