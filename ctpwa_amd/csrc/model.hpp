// Host-side model for the native likelihood path.
//
// The plan text (written by ctpwa_amd/model.py) describes the decay topologies, their
// (S, L) waves, the resonance blocks, the amplitude order and the free-parameter
// slots exactly as the reference enumerates them (src/Info.cu:43-357,
// src/Analysis.cu:4481-4830, src/AmpGen.cu:2009-2242).  This header parses it and
// derives the device-side layout:
//
//  * kinematic nodes   - distinct (mother, daughter1, daughter2) final-state subsets;
//  * invariants        - event-only scalars the lineshape denominators need
//                        (s = m^2, the BWR width factor q^(2L+1)/(m N_L(qd)), Flatte
//                        phase-space factors, Gounaris-Sakurai terms);
//  * SL tensors ("u")  - per (topology, SL combination, barrier mask): the spin
//                        amplitude with the event-only part of every Blatt-Weisskopf
//                        factor folded in;
//  * denominators      - one per resonance propagator: D = sum_t c_t x_t - i sum_t d_t y_t,
//                        with theta-only coefficients c_t, d_t and event invariants x, y;
//  * amplitude numerators K_a(theta) (real) - the theta-only barrier parts.
//
// Every lineshape of the reference's built-in set factorises this way:
// Bf(L,q,q0) = sqrt(N_L(q0 d)) / sqrt(N_L(q d)) (include/ResModel.cuh:184-211) and
// m0 Gamma(m) of the BWR (include/ResModel.cuh:213-223) is a theta-only factor times
// q^(2L+1) / (m N_L(q d)).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace ctpwa {

enum Model { M_BWR = 0, M_BW = 1, M_ONE = 2, M_FLATTE = 3, M_GS = 4, M_CUSTOM = 5, M_HIST = 6, M_INTERP = 7 };
enum InvKind { INV_ONE = 0, INV_S = 1, INV_E = 2, INV_RHO_RE = 3, INV_RHO_IM = 4,
               INV_GS_Q2H = 5, INV_GS_Q2 = 6, INV_GS_Q3M = 7, INV_M = 8, INV_Q = 9,
               INV_TAB_RE = 10, INV_TAB_IM = 11,     // table lineshape value at m (L: table offset, d: model)
               INV_DAUGHTER = 12 };                  // 12..19: |p|, E, cos(theta), phi of daughter 1, then 2
                                                     // (event frame; Custom expressions)

struct PNode { int dj, dj1, dj2, d1ref, d2ref, has_bf, lmin; double bf_d, mm, md1, md2; };
struct PTopo { std::vector<PNode> nodes; std::vector<std::vector<std::pair<int, int>>> slc; };
struct PRes {
    int node, model;
    std::vector<double> params;
    std::vector<int> slots;
    std::vector<double> ch;
    std::vector<std::string> expr;   // Custom: expression tokens in prefix order
    // Hist: [m_min, m_max, n, values...]; Interp: [method, N, x_min, dx, re..., im...]
    // (equal spacing) or [-(method+1), N, x..., re..., im...]
    std::vector<double> tab;
};
// c0: the first of the topology's SL combinations that the block's waves use (a block
// split into runs of one L for a Custom expression that takes each wave's L)
struct PBlock { int topo, site, nsl, c0 = 0; std::vector<PRes> res; };

struct Plan {
    int nfinal = 0, npol = 0;
    std::vector<int> polmap;
    std::vector<PTopo> topos;
    std::vector<PBlock> blocks;
    int ntheta = 0, namp = 0, nfree = 0;
    std::vector<double> th_init, th_lo, th_hi;
    std::vector<int> amap;
    std::vector<double> aratio;
    std::vector<int> g_slot;
    std::vector<double> g_mu, g_sig;
    // partial wave of every block (a block and its identical-particle permuted copies
    // form one wave); waves are numbered 0 .. nwave-1
    std::vector<int> bwave;
    int nwave = 0;
};

inline Plan parse_plan(const std::string& text) {
    Plan P;
    std::istringstream in(text);
    std::string line;
    if (!std::getline(in, line) || line.rfind("CTPWA_PLAN", 0) != 0)
        throw std::runtime_error("plan text does not start with CTPWA_PLAN");
    bool ended = false;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key)) continue;
        if (key == "nfinal") ls >> P.nfinal;
        else if (key == "npol") { ls >> P.npol; P.polmap.resize(P.npol); for (auto& x : P.polmap) ls >> x; }
        else if (key == "ntopo" || key == "nblock") {}
        else if (key == "topo") { P.topos.emplace_back(); }
        else if (key == "node") {
            if (P.topos.empty()) throw std::runtime_error("node before topo");
            PNode n; int k;
            ls >> k >> n.dj >> n.dj1 >> n.dj2 >> n.d1ref >> n.d2ref >> n.has_bf >> n.bf_d >> n.lmin >> n.mm >> n.md1 >> n.md2;
            P.topos.back().nodes.push_back(n);
        } else if (key == "slc") {
            std::vector<std::pair<int, int>> c; int S, L;
            while (ls >> S >> L) c.push_back({S, L});
            P.topos.back().slc.push_back(c);
        } else if (key == "block") {
            PBlock b; int idx, nres;
            ls >> idx >> b.topo >> b.site >> b.nsl >> nres;
            P.blocks.push_back(b);
        } else if (key == "res") {
            if (P.blocks.empty()) throw std::runtime_error("res before block");
            PRes r; int np, nch;
            ls >> r.node >> r.model >> np;
            r.params.resize(np); for (auto& x : r.params) ls >> x;
            r.slots.resize(np); for (auto& x : r.slots) ls >> x;
            ls >> nch; r.ch.resize(2 * nch); for (auto& x : r.ch) ls >> x;
            P.blocks.back().res.push_back(r);
        } else if (key == "cexpr") {
            if (P.blocks.empty() || P.blocks.back().res.empty()) throw std::runtime_error("cexpr before res");
            int n; ls >> n;
            if (n <= 0 || n > 100000) throw std::runtime_error("bad cexpr length");
            std::vector<std::string> toks(n);
            for (auto& t : toks) if (!(ls >> t)) throw std::runtime_error("cexpr is truncated");
            P.blocks.back().res.back().expr = toks;
        } else if (key == "rtab") {
            if (P.blocks.empty() || P.blocks.back().res.empty()) throw std::runtime_error("rtab before res");
            long long n; ls >> n;
            if (n <= 0 || n > 10000000) throw std::runtime_error("bad rtab length");
            std::vector<double> t((size_t)n);
            for (auto& x : t) if (!(ls >> x)) throw std::runtime_error("rtab is truncated");
            P.blocks.back().res.back().tab = t;
        } else if (key == "ntheta") ls >> P.ntheta;
        else if (key == "theta") { int j; double a, lo, hi; ls >> j >> a >> lo >> hi; P.th_init.push_back(a); P.th_lo.push_back(lo); P.th_hi.push_back(hi); }
        else if (key == "namp") ls >> P.namp >> P.nfree;
        else if (key == "amap") { P.amap.resize(P.namp); for (auto& x : P.amap) ls >> x; }
        else if (key == "aratio") { P.aratio.resize(P.namp); for (auto& x : P.aratio) ls >> x; }
        else if (key == "bslc0") {
            for (auto& B : P.blocks)
                if (!(ls >> B.c0)) throw std::runtime_error("plan: bslc0 is truncated");
        }
        else if (key == "bwave") {
            P.bwave.resize(P.blocks.size());
            for (auto& x : P.bwave)
                if (!(ls >> x)) throw std::runtime_error("plan: bwave is truncated");
        }
        else if (key == "ngauss") {}
        else if (key == "gauss") { int s; double mu, sg; ls >> s >> mu >> sg; P.g_slot.push_back(s); P.g_mu.push_back(mu); P.g_sig.push_back(sg); }
        else if (key == "end") { ended = true; break; }
        else throw std::runtime_error("unknown plan record '" + key + "'");
    }
    if (!ended) throw std::runtime_error("plan text is truncated (no 'end')");
    if (P.nfinal <= 0 || P.nfinal > 16) throw std::runtime_error("plan: nfinal must be in [1, 16]");
    if (P.npol <= 0) throw std::runtime_error("plan: npol must be positive");
    if ((int)P.th_init.size() != P.ntheta) throw std::runtime_error("plan: theta count mismatch");
    for (const PBlock& B : P.blocks) {
        if (B.topo < 0 || B.topo >= (int)P.topos.size()) throw std::runtime_error("plan: block topology out of range");
        if (B.nsl < 0 || B.c0 < 0 || (size_t)B.c0 + B.nsl > P.topos[B.topo].slc.size())
            throw std::runtime_error("plan: a block's SL combinations exceed its topology's");
    }
    if (P.bwave.empty())
        for (size_t b = 0; b < P.blocks.size(); ++b) P.bwave.push_back((int)b);
    if (P.bwave.size() != P.blocks.size()) throw std::runtime_error("plan: bwave count mismatch");
    for (int w : P.bwave) P.nwave = std::max(P.nwave, w + 1);
    std::vector<int> seen(P.nwave, 0);
    for (int w : P.bwave) {
        if (w < 0) throw std::runtime_error("plan: negative partial-wave index");
        seen[w] = 1;
    }
    for (int x : seen)
        if (!x) throw std::runtime_error("plan: partial-wave indices are not contiguous");
    return P;
}

// ---------------------------------------------------------------------------
// Derived layout
// ---------------------------------------------------------------------------
struct KinNode { uint32_t mset, d1set, d2set; };

struct InvSpec {
    int kin, kind, L;
    double d, ma, mb;
    bool operator<(const InvSpec& o) const {
        return std::tie(kin, kind, L, d, ma, mb) < std::tie(o.kin, o.kind, o.L, o.d, o.ma, o.mb);
    }
};

struct USpec { int topo, slc; uint32_t mask; };   // mask bit k: fold node k's barrier

struct Term { int inv; int coef; double cval; };   // inv < 0: constant 1; coef < 0: constant cval
struct DenSpec {
    int block, res, node;
    // D = sum re - i sum im; a static factor (table lineshapes) is instead the factor
    // F = sum re + i sum im itself, without theta dependence
    std::vector<Term> re, im;
    bool stat = false;
    int ncoef = 0;              // theta-dependent coefficients (slots in the table)
    int coef_base = 0;          // index of the first coefficient record
};

struct AmpSpec { int block, slc, u; int coef_base; };

// one accumulator of the theta gradient: sum_e x_t(e) Re R_d(e) (real term) or
// sum_e y_t(e) Im R_d(e) (imaginary term) of a theta-dependent denominator coefficient
struct ThTerm { int den, coef, im, inv; };

// one Custom lineshape factor of a block: evaluated per event with forward-mode
// derivatives over the block's local variables
struct CustomSpec {
    int block, res, node;
    int inv_m, inv_q;                  // invariant rows of m and q at the node
    int inv_d[8];                      // rows of the daughter-momentum variables (-1: unused)
    std::vector<std::string> expr;
    std::vector<int> param_slot;       // parameter k -> theta slot (-1: fixed value)
    std::vector<double> param_val;
};
// one accumulator of the theta gradient through a Custom factor: sum_e Re(z F'_j other)
struct CTerm { int custom, local; };

struct Model_ {
    Plan P;
    std::vector<KinNode> kin;
    std::vector<std::vector<int>> topo_kin;      // [topo][node] -> kin index
    std::vector<InvSpec> inv;
    std::vector<USpec> us;
    std::vector<DenSpec> dens;
    std::vector<std::vector<int>> block_dens;     // [block] -> den indices
    std::vector<AmpSpec> amps;                    // [namp]
    std::vector<std::vector<int>> bvars;          // [block] -> local var -> global slot
    std::vector<ThTerm> thterms;                  // theta-gradient accumulators
    std::vector<CustomSpec> customs;
    std::vector<std::vector<int>> block_customs;  // [block] -> custom indices
    std::vector<CTerm> cterms;                    // Custom theta-gradient accumulators
    std::vector<double> tabs;                     // table lineshapes, concatenated
    int maxloc = 0;
    // code-generation options of the fused kernel
    int min_waves = 0;       // __launch_bounds__ minimum waves per SIMD (0: compiler choice)
    int wgs = 128;           // threads (= events per tile) of a workgroup: 64, 128 or 256
    bool lds_pipe = false;   // copy the next tile's inputs HBM -> LDS while computing the current one
    int prefetch_sl = 0;     // load the SL values and invariants of the next 1 or 2 events ahead (registers)
    bool inv_tree = false;   // simultaneous inversion over a balanced product tree (else a prefix chain)
    // storage precision of the per-event inputs: 0 FP64; 1 mixed (FP32 SL tensors and
    // invariants, FP64 squared masses and Custom-expression m, q)
    int mixed = 0;
    // polarisations stored per SL tensor: npol, or 1 when the two polarisations of every SL
    // tensor are mirror images, sl_u,1 = psign[u] conj(sl_u,0) (set at handle creation
    // after a check over all events; spin-0 final states of a J = M, -M initial state)
    int nps = 0;
    std::vector<int> psign;
    int npstore() const { return nps ? nps : P.npol; }
    // storage row of invariant i: >= 0 row of the FP64 store, < 0: -(row of the FP32 store) - 1
    std::vector<int> imap() const {
        std::vector<int> m(inv.size());
        int nd = 0, nf = 0;
        for (size_t i = 0; i < inv.size(); ++i) {
            // squared masses, and the m and q of Custom expressions (arbitrary formulas
            // that may cancel like m0^2 - m^2), stay FP64
            if (!mixed || inv[i].kind == INV_S || inv[i].kind == INV_M || inv[i].kind == INV_Q ||
                inv[i].kind == INV_TAB_RE || inv[i].kind == INV_TAB_IM || inv[i].kind >= INV_DAUGHTER) m[i] = nd++;
            else m[i] = -(nf++) - 1;
        }
        return m;
    }
    // coefficient records: each holds value, NLOC first and NLOC*NLOC second derivatives
    static const int NLOC = 8;
    static const int REC = 1 + NLOC + NLOC * NLOC;
    int n_coef_rec = 0;                           // amplitude K records + denominator coefficient records
    int amp_rec0 = 0;                             // first amplitude K record
    // table layout (doubles): [v: 2 namp][c: 2 namp][records]
    size_t off_v() const { return 0; }
    size_t off_c() const { return 2 * (size_t)P.namp; }
    size_t off_rec() const { return 4 * (size_t)P.namp; }
    size_t table_size() const { return off_rec() + (size_t)n_coef_rec * REC + 2 * (size_t)P.ntheta; }
    int nacc() const { return 1 + 2 * P.namp + (int)thterms.size() + (int)cterms.size(); }
    // Custom factors: derivative directions (the largest block among Custom blocks) and
    // per-event rows of one factor in the factor buffer (value, then dr[], di[])
    int ncv() const {
        size_t n = 1;
        for (const CustomSpec& C : customs) n = std::max(n, bvars[C.block].size());
        return (int)n;
    }
    int ncf() const { return 2 + 2 * ncv(); }
    // the Hessian's second-order rows of one factor: value, dr[], di[], hr[], hi[] (packed k <= l)
    int ncf2() const { return 2 + 2 * ncv() + ncv() * (ncv() + 1); }
    size_t off_th() const { return off_rec() + (size_t)n_coef_rec * REC; }   // theta values (for Custom)
    size_t off_lth() const { return off_th() + (size_t)P.ntheta; }          // theta of the last NLL / Hessian
};

// final-state subset of a momentum reference
inline uint32_t ref_set(const PTopo& T, int ref, int nfinal, int depth = 0) {
    if (depth > 64) throw std::runtime_error("cyclic decay topology");
    if (ref >= 0) return 1u << ref;
    int k = -ref - 1;
    if (k <= 0 || k >= (int)T.nodes.size()) throw std::runtime_error("bad node reference");
    return ref_set(T, T.nodes[k].d1ref, nfinal, depth + 1) | ref_set(T, T.nodes[k].d2ref, nfinal, depth + 1);
}

inline Model_ derive(const Plan& P) {
    Model_ M;
    M.P = P;
    uint32_t all = (P.nfinal >= 32) ? 0xffffffffu : ((1u << P.nfinal) - 1u);
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, int> kin_idx;
    for (const PTopo& T : P.topos) {
        std::vector<int> tk;
        for (size_t k = 0; k < T.nodes.size(); ++k) {
            const PNode& n = T.nodes[k];
            uint32_t d1 = ref_set(T, n.d1ref, P.nfinal), d2 = ref_set(T, n.d2ref, P.nfinal);
            uint32_t m = (k == 0) ? all : (d1 | d2);
            auto key = std::make_tuple(m, d1, d2);
            auto it = kin_idx.find(key);
            if (it == kin_idx.end()) {
                kin_idx[key] = (int)M.kin.size();
                tk.push_back((int)M.kin.size());
                M.kin.push_back(KinNode{m, d1, d2});
            } else tk.push_back(it->second);
            if (n.dj * n.dj1 * n.dj2 > 64) throw std::runtime_error("node tensor larger than 64 entries");
        }
        if (T.nodes.size() > 8) throw std::runtime_error("more than 8 decay steps in a chain");
        // contracted-tensor sizes (labels: mother replaced by its two daughters, in order)
        {
            std::vector<int> shape = {T.nodes[0].dj, T.nodes[0].dj1, T.nodes[0].dj2};
            std::vector<int> lab = {-1, T.nodes[0].d1ref, T.nodes[0].d2ref};
            for (size_t k = 1; k < T.nodes.size(); ++k) {
                int ci = -1;
                for (size_t j = 0; j < lab.size(); ++j) if (lab[j] == -(int)k - 1) { ci = (int)j; break; }
                if (ci < 0) throw std::runtime_error("decay step whose mother is not produced earlier in the chain");
                shape.erase(shape.begin() + ci);
                lab.erase(lab.begin() + ci);
                shape.push_back(T.nodes[k].dj1); lab.push_back(T.nodes[k].d1ref);
                shape.push_back(T.nodes[k].dj2); lab.push_back(T.nodes[k].d2ref);
                int sz = 1;
                for (int d : shape) sz *= d;
                if (sz > 64) throw std::runtime_error("contracted spin tensor larger than 64 entries");
            }
            int sz = 1;
            for (int d : shape) sz *= d;
            for (int pm : P.polmap)
                if (pm < 0 || pm >= sz) throw std::runtime_error("polarisation map index outside the spin tensor");
        }
        M.topo_kin.push_back(tk);
    }
    std::map<InvSpec, int> inv_idx;
    auto add_inv = [&](InvSpec s) -> int {
        if (s.kind == INV_ONE) return -1;          // the constant 1 needs no storage
        auto it = inv_idx.find(s);
        if (it != inv_idx.end()) return it->second;
        int i = (int)M.inv.size();
        inv_idx[s] = i;
        M.inv.push_back(s);
        return i;
    };
    std::map<std::tuple<int, int, uint32_t>, int> u_idx;
    std::vector<std::pair<int, size_t>> tab_starts;   // offset and length of each table in M.tabs
    int rec = 0;
    M.block_dens.resize(P.blocks.size());
    M.block_customs.resize(P.blocks.size());
    M.amps.resize(P.namp);
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const PBlock& B = P.blocks[b];
        const PTopo& T = P.topos[B.topo];
        int nn = (int)T.nodes.size();
        // local variables of the block
        std::vector<int> vars;
        for (const PRes& r : B.res)
            for (int s : r.slots)
                if (s >= 0 && std::find(vars.begin(), vars.end(), s) == vars.end()) vars.push_back(s);
        if ((int)vars.size() > Model_::NLOC) throw std::runtime_error("more than 8 free parameters in one block");
        M.maxloc = std::max(M.maxloc, (int)vars.size());
        M.bvars.push_back(vars);
        std::vector<int> res_at(nn, -1);
        for (size_t r = 0; r < B.res.size(); ++r) {
            int k = B.res[r].node;
            if (k <= 0 || k >= nn) throw std::runtime_error("resonance is not attached to a decaying intermediate");
            res_at[k] = (int)r;
        }
        // barrier mask of this block
        uint32_t mask = 0;
        for (int k = 0; k < nn; ++k) {
            bool fold;
            if (k == 0 || res_at[k] < 0) fold = T.nodes[k].has_bf;
            else {
                int mdl = B.res[res_at[k]].model;
                fold = T.nodes[k].has_bf && mdl != M_BW && mdl != M_CUSTOM && mdl != M_HIST && mdl != M_INTERP;
            }
            if (fold) mask |= 1u << k;
        }
        // denominators (a block without waves contributes nothing but keeps its slots)
        for (size_t r = 0; r < B.res.size() && B.nsl > 0; ++r) {
            const PRes& R = B.res[r];
            int k = R.node;
            int kin = M.topo_kin[B.topo][k];
            const PNode& n = T.nodes[k];
            DenSpec D{(int)b, (int)r, k, {}, {}, 0, rec};
            auto S = add_inv({kin, INV_S, 0, 0.0, 0.0, 0.0});
            auto ONE = add_inv({kin, INV_ONE, 0, 0.0, 0.0, 0.0});
            switch (R.model) {
            case M_ONE: continue;
            case M_CUSTOM: {
                CustomSpec C;
                C.block = (int)b; C.res = (int)r; C.node = k;
                C.inv_m = add_inv({kin, INV_M, 0, 0.0, 0.0, 0.0});
                C.inv_q = add_inv({kin, INV_Q, 0, 0.0, 0.0, 0.0});
                C.expr = R.expr;
                static const char* dvar[8] = {"v:p1p", "v:p1e", "v:p1costheta", "v:p1phi",
                                              "v:p2p", "v:p2e", "v:p2costheta", "v:p2phi"};
                for (int j = 0; j < 8; ++j) {
                    C.inv_d[j] = -1;
                    if (std::find(C.expr.begin(), C.expr.end(), std::string(dvar[j])) != C.expr.end())
                        C.inv_d[j] = add_inv({kin, INV_DAUGHTER + j, 0, 0.0, 0.0, 0.0});
                }
                if (C.expr.empty()) throw std::runtime_error("Custom lineshape without expression");
                C.param_slot = R.slots;
                C.param_val = R.params;
                M.block_customs[b].push_back((int)M.customs.size());
                M.customs.push_back(C);
                continue;
            }
            case M_BWR: {
                if (R.params.size() < 2) throw std::runtime_error("BWR needs mass and width");
                int E = add_inv({kin, INV_E, n.lmin, n.has_bf ? n.bf_d : 0.0, 0.0, 0.0});
                D.re = {{ONE, 0, 0.0}, {S, -1, -1.0}};
                D.im = {{E, 1, 0.0}};
                D.ncoef = 2;
                break;
            }
            case M_BW: {
                if (R.params.size() < 2) throw std::runtime_error("BW needs mass and width");
                D.re = {{ONE, 0, 0.0}, {S, -1, -1.0}};
                D.im = {{ONE, 1, 0.0}};
                D.ncoef = 2;
                break;
            }
            case M_FLATTE: {
                int nch = (int)R.ch.size() / 2;
                if ((int)R.params.size() < 1 + nch) throw std::runtime_error("Flatte needs one coupling per channel");
                D.re = {{ONE, 0, 0.0}, {S, -1, -1.0}};
                for (int i = 0; i < nch; ++i) {
                    int ri = add_inv({kin, INV_RHO_IM, 0, 0.0, R.ch[2 * i], R.ch[2 * i + 1]});
                    int rr = add_inv({kin, INV_RHO_RE, 0, 0.0, R.ch[2 * i], R.ch[2 * i + 1]});
                    D.re.push_back({ri, 1 + i, 0.0});
                    D.im.push_back({rr, 1 + i, 0.0});
                }
                D.ncoef = 1 + nch;
                break;
            }
            case M_HIST:
            case M_INTERP: {
                // value of the table at the pair mass, no vertex barrier
                // (include/ResModel.cuh:67-85,153-157,226-290)
                // the device reads the table through its header: check every length here
                const std::vector<double>& tb = R.tab;
                if (R.model == M_HIST) {
                    if (tb.size() < 5 || tb[2] != std::floor(tb[2]) || tb[2] < 2 || tb.size() != 3 + (size_t)tb[2] ||
                        !(tb[1] > tb[0]))
                        throw std::runtime_error("Hist table must be [m_min, m_max, n >= 2, n values] with m_max > m_min");
                } else {
                    if (tb.size() < 2 || tb[1] != std::floor(tb[1]) || tb[1] < 2)
                        throw std::runtime_error("Interp table needs at least two points");
                    size_t N = (size_t)tb[1];
                    bool uniform = tb[0] >= 0;
                    if ((uniform && (tb[0] != 1.0 || tb.size() != 4 + 2 * N || !(tb[3] != 0.0))) ||
                        (!uniform && (tb[0] != -2.0 || tb.size() != 2 + 3 * N)))
                        throw std::runtime_error("Interp table must be [1, N, x_min, dx, re, im] or [-2, N, x, re, im]");
                }
                int off = -1;
                for (const auto& ts : tab_starts)
                    if (ts.second == R.tab.size() && std::equal(R.tab.begin(), R.tab.end(), M.tabs.begin() + ts.first)) {
                        off = ts.first;
                        break;
                    }
                if (off < 0) {
                    off = (int)M.tabs.size();
                    tab_starts.push_back({off, R.tab.size()});
                    M.tabs.insert(M.tabs.end(), R.tab.begin(), R.tab.end());
                }
                int tre = add_inv({kin, INV_TAB_RE, off, (double)R.model, 0.0, 0.0});
                D.re = {{tre, -1, 1.0}};
                if (R.model == M_INTERP) D.im = {{add_inv({kin, INV_TAB_IM, off, (double)R.model, 0.0, 0.0}), -1, 1.0}};
                D.ncoef = 0;
                D.stat = true;
                break;
            }
            case M_GS: {
                if (R.params.size() < 2) throw std::runtime_error("GS needs mass and width");
                double mpi = R.ch.empty() ? 0.1396 : R.ch[0];
                int Q2H = add_inv({kin, INV_GS_Q2H, 0, 0.0, mpi, 0.0});
                int Q2 = add_inv({kin, INV_GS_Q2, 0, 0.0, 0.0, 0.0});
                int Q3M = add_inv({kin, INV_GS_Q3M, 0, 0.0, 0.0, 0.0});
                // D = A0 - C s + c1 q^2 h - c2 q^2 - i c1 q^3/m
                D.re = {{ONE, 0, 0.0}, {S, 1, 0.0}, {Q2H, 2, 0.0}, {Q2, 3, 0.0}};
                D.im = {{Q3M, 2, 0.0}};
                D.ncoef = 4;
                break;
            }
            default: throw std::runtime_error("lineshape model not supported");
            }
            D.coef_base = rec;
            rec += D.ncoef;
            M.block_dens[b].push_back((int)M.dens.size());
            M.dens.push_back(D);
        }
        // amplitudes
        for (int c = 0; c < B.nsl; ++c) {
            const int sc = B.c0 + c;
            auto key = std::make_tuple(B.topo, sc, mask);
            int u;
            auto it = u_idx.find(key);
            if (it == u_idx.end()) {
                u = (int)M.us.size();
                u_idx[key] = u;
                M.us.push_back(USpec{B.topo, sc, mask});
            } else u = it->second;
            int a = B.site + c;
            if (a < 0 || a >= P.namp) throw std::runtime_error("amplitude index out of range");
            M.amps[a] = AmpSpec{(int)b, sc, u, 0};
        }
    }
    for (size_t d = 0; d < M.dens.size(); ++d) {
        const DenSpec& Dn = M.dens[d];
        if (M.bvars[Dn.block].empty()) continue;
        for (const Term& t : Dn.re) if (t.coef >= 0) M.thterms.push_back(ThTerm{(int)d, t.coef, 0, t.inv});
        for (const Term& t : Dn.im) if (t.coef >= 0) M.thterms.push_back(ThTerm{(int)d, t.coef, 1, t.inv});
    }
    for (size_t c = 0; c < M.customs.size(); ++c) {
        const std::vector<int>& vars = M.bvars[M.customs[c].block];
        for (size_t li = 0; li < vars.size(); ++li) {
            bool uses = false;
            for (int s : M.customs[c].param_slot) if (s == vars[li]) uses = true;
            if (uses) M.cterms.push_back(CTerm{(int)c, (int)li});
        }
    }
    // amplitude records follow the denominator records
    M.amp_rec0 = rec;
    for (int a = 0; a < P.namp; ++a) M.amps[a].coef_base = rec + a;
    M.n_coef_rec = rec + P.namp;
    if (P.topos.empty() || P.namp <= 0) throw std::runtime_error("plan has no amplitudes");
    return M;
}

}  // namespace ctpwa
