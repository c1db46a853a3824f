// AXON CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Each function restates one reference kernel or host routine of src/axon.mm;
// the citation is on the function.  The column-parallel mode splits columns
// over std::thread workers; every cross-column quantity is an integer sum, so
// the parallel mode is bit-identical to the single-threaded one.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t NPC = 32, NSLOT = 32, SLOT_U4 = 2, PRESYN = 256, FANIN = 8192;
constexpr uint32_t COLS_PER_GRP = 8;

// mm:54 (lowbias32)
inline uint32_t hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
inline uint32_t popc(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

// Params, mm:56-83 / mm:650-652
struct Params {
    uint32_t k, theta, dthr, boost, mode, inCols, step, learn, depress, targetCols, clamp,
        freeze, biasUp, biasDown, biasShift, histOn, useBias, phase, bandLo, bandHi,
        marginGate, hetero, ltdRep, biasDownEvery, facil, pad;
};

// ---- bit-serial counter arithmetic on one 32-synapse word, mm:371-420 ----
inline void bs_inc(uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3, uint32_t m) {
    uint32_t c = m, t;
    t = a0 ^ c; c = a0 & c; a0 = t;
    t = a1 ^ c; c = a1 & c; a1 = t;
    t = a2 ^ c; c = a2 & c; a2 = t;
    t = a3 ^ c; c = a3 & c; a3 = t;
    a0 |= c; a1 |= c; a2 |= c; a3 |= c;
}
inline void bs_dec(uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3, uint32_t m) {
    uint32_t b = m, t;
    t = a0 ^ b; b = (~a0) & b; a0 = t;
    t = a1 ^ b; b = (~a1) & b; a1 = t;
    t = a2 ^ b; b = (~a2) & b; a2 = t;
    t = a3 ^ b; b = (~a3) & b; a3 = t;
    uint32_t keep = ~b; a0 &= keep; a1 &= keep; a2 &= keep; a3 &= keep;
}
inline uint32_t bs_lt(uint32_t a3, uint32_t a2, uint32_t a1, uint32_t a0, uint32_t K) {
    uint32_t lt = 0, eq = 0xFFFFFFFFu, kb;
    kb = (K >> 3) & 1u; if (kb) lt |= eq & ~a3; eq &= kb ? a3 : ~a3;
    kb = (K >> 2) & 1u; if (kb) lt |= eq & ~a2; eq &= kb ? a2 : ~a2;
    kb = (K >> 1) & 1u; if (kb) lt |= eq & ~a1; eq &= kb ? a1 : ~a1;
    kb = (K >> 0) & 1u; if (kb) lt |= eq & ~a0;
    return lt;
}
inline uint32_t bs_gt(uint32_t a3, uint32_t a2, uint32_t a1, uint32_t a0, uint32_t K) {
    uint32_t gt = 0, eq = 0xFFFFFFFFu, kb;
    kb = (K >> 3) & 1u; if (!kb) gt |= eq & a3; eq &= kb ? a3 : ~a3;
    kb = (K >> 2) & 1u; if (!kb) gt |= eq & a2; eq &= kb ? a2 : ~a2;
    kb = (K >> 1) & 1u; if (!kb) gt |= eq & a1; eq &= kb ? a1 : ~a1;
    kb = (K >> 0) & 1u; if (!kb) gt |= eq & a0;
    return gt;
}
inline void bs_set(uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3, uint32_t m, uint32_t C) {
    uint32_t keep = ~m;
    a0 = (a0 & keep) | (((C >> 0) & 1u) ? m : 0u);
    a1 = (a1 & keep) | (((C >> 1) & 1u) ? m : 0u);
    a2 = (a2 & keep) | (((C >> 2) & 1u) ? m : 0u);
    a3 = (a3 & keep) | (((C >> 3) & 1u) ? m : 0u);
}
inline void bs_clip(uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3, uint32_t lo, uint32_t hi) {
    uint32_t l = bs_lt(a3, a2, a1, a0, lo);
    uint32_t g = bs_gt(a3, a2, a1, a0, hi);
    bs_set(a0, a1, a2, a3, l, lo);
    bs_set(a0, a1, a2, a3, g, hi);
}

// kwta_ballot_margin, mm:134-144: rank_l = #{j : v_j > v_l or (v_j == v_l and j < l)}
uint32_t kwta_margin(const uint32_t v[32], uint32_t k, uint32_t& vk, uint32_t& vk1) {
    uint32_t rank[32];
    for (uint32_t l = 0; l < 32; ++l) {
        uint32_t r = 0;
        for (uint32_t j = 0; j < 32; ++j) r += (v[j] > v[l] || (v[j] == v[l] && j < l)) ? 1u : 0u;
        rank[l] = r;
    }
    vk = 0; vk1 = 0;
    uint32_t ball = 0;
    for (uint32_t l = 0; l < 32; ++l) {
        vk = std::max(vk, rank[l] == k - 1u ? v[l] : 0u);
        vk1 = std::max(vk1, rank[l] == k ? v[l] : 0u);
        if (rank[l] < k) ball |= 1u << l;
    }
    return ball;
}

// k_topk_par (mm:301-352) and k_topk_serial (mm:259-292).  They agree whenever
// a crossing bin exists; the degenerate paths are restated literally.
void topk(const uint32_t* drive, uint32_t nCol, uint32_t inCols, uint32_t target, bool par,
          uint32_t theta[2]) {
    uint32_t need = target, prefix = 0;
    for (int by = 2; by >= 0; --by) {
        uint32_t hist[256] = {0};
        uint32_t shHi = (uint32_t)(by + 1) * 8u, shLo = (uint32_t)by * 8u;
        for (uint32_t c = inCols; c < nCol; ++c) {
            uint32_t d = drive[c];
            if ((d >> shHi) == (prefix >> shHi)) hist[(d >> shLo) & 255u]++;
        }
        if (par) {
            uint32_t scan[257];
            scan[256] = 0;
            for (int i = 255; i >= 0; --i) scan[i] = scan[i + 1] + hist[i];
            uint32_t selBin = 0;
            for (uint32_t t = 0; t < 256; ++t) {
                uint32_t above = (t + 1u < 256u) ? scan[t + 1] : 0u;
                if (scan[t] >= need && above < need) selBin = t;
            }
            uint32_t above = (selBin + 1u < 256u) ? scan[selBin + 1] : 0u;
            need = need - above;
            prefix |= (selBin << shLo);
        } else {
            uint32_t cum = 0, b = 0, above = 0;
            for (int i = 255; i >= 0; --i) {
                uint32_t h = hist[i];
                if (cum + h >= need) { b = (uint32_t)i; above = cum; break; }
                cum += h;
            }
            need = need - above;
            prefix |= (b << shLo);
        }
    }
    theta[0] = std::max(prefix, 1u);
    theta[1] = need;
}

// libc++ std::shuffle (call site mm:797) with uniform_int_distribution<ptrdiff_t>:
// Fisher-Yates from the front; j in [0, n-1-i] drawn by masked rejection over
// 32-bit mt19937 words (w = ceil(log2(range)) low bits of one word per try).
template <class G>
uint64_t libcxx_uid(G& g, uint64_t rp) {
    if (rp == 1) return 0;
    uint32_t w = 64u - (uint32_t)__builtin_clzll(rp) - 1u;
    if ((rp & (~0ull >> (64u - w))) != 0) ++w;
    if (w > 32) { fprintf(stderr, "oracle: shuffle range too large\n"); abort(); }
    uint32_t mask = w ? (0xFFFFFFFFu >> (32u - w)) : 0u;
    uint64_t u;
    do { u = (uint64_t)((uint32_t)g() & mask); } while (u >= rp);
    return u;
}
template <class G>
void libcxx_shuffle(uint32_t* a, uint32_t n, G& g) {
    int64_t d = n;
    if (d > 1) {
        int64_t first = 0, last = d - 1;
        --d;
        for (; first < last; ++first, --d) {
            int64_t i = (int64_t)libcxx_uid(g, (uint64_t)d + 1u);
            if (i != 0) std::swap(a[first], a[first + i]);
        }
    }
}
// uniform_real_distribution<double>(0,1) over mt19937: two words, low first.
template <class G>
double uniform01(G& g) {
    double s = (double)(uint32_t)g();
    double base = 4294967296.0;
    s += (double)(uint32_t)g() * base;
    base *= 4294967296.0;
    return s / base;
}

template <class F>
void par_cols(int nthreads, uint32_t c0, uint32_t c1, F&& f) {
    uint32_t n = c1 > c0 ? c1 - c0 : 0;
    if (nthreads <= 1 || n < 64) { f(c0, c1, 0); return; }
    std::vector<std::thread> th;
    uint32_t per = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        uint32_t a = c0 + t * per, b = std::min(c1, a + per);
        if (a >= b) break;
        th.emplace_back([&, a, b, t] { f(a, b, t); });
    }
    for (auto& x : th) x.join();
}

}  // namespace

struct ora_net {
    ora_config cfg;
    uint32_t NCOL, N, NGRP, IN_COLS, MAXT, NSTATS, NDLOG;
    size_t PLANE_U4;
    std::vector<uint32_t> W[4];
    std::vector<uint32_t> conn, spkA, spkB, drv, bbits, acc, bias, marg, refBall, driveLog,
        mult, stim, stat, hist, theta;
    std::vector<uint64_t> cntProp, cntLearn;  // per column: tiles read / tiles rewritten
    std::vector<std::vector<uint32_t>> stimCols;
    Params P;
    uint32_t K, kTest, facil, contrast, ltdEvery, topkPar;
    int nthreads;
};

// ------------------------------------------------------------------ config
extern "C" void ora_config_defaults(ora_config* c) {
    memset(c, 0, sizeof(*c));
    // built-in defaults, mm:688-728 (NOT Config C)
    c->nand = 4; c->target = 256; c->k = 6; c->dthr = 2; c->boost = 4; c->mode = 1;
    c->epochs = 30; c->T = 10; c->clampsteps = 5; c->stim = 12; c->quiet = 0; c->seed = 0;
    c->split = 0; c->unionchar = 0; c->biasup = 0; c->biasdown = 1; c->biasshift = 4;
    c->biasdownevery = 1; c->inbr = 8; c->stimcols = 128; c->incols = 512; c->ncol = 4096;
    c->diag = 0; c->profile = 0; c->topkpar = 1; c->vtopk = 0; c->contrast = 0; c->bandlo = 5;
    c->bandhi = 11; c->margingate = 1; c->hetero = 0; c->vbs = 0; c->ltdrep = 1; c->ltdevery = 1;
    c->ktest = 0; c->facil = 0; c->deep = 0; c->wave = 0; c->wavet = 35;
    c->trainclamp = c->contrast ? c->clampsteps : c->T;
    c->max_t = 64;
}

// argu(), mm:660-667
static uint32_t argu(int argc, const char* const* argv, const char* key, uint32_t dflt) {
    size_t kl = strlen(key);
    uint32_t v = dflt;
    for (int i = 1; i < argc; ++i)
        if (!strncmp(argv[i], key, kl) && argv[i][kl] == '=') v = (uint32_t)atoi(argv[i] + kl + 1);
    return v;
}

extern "C" void ora_config_parse(int argc, const char* const* argv, ora_config* c) {
    ora_config_defaults(c);
    c->nand = argu(argc, argv, "nand", 4);
    c->target = argu(argc, argv, "target", 256);
    c->k = argu(argc, argv, "k", 6);
    c->dthr = argu(argc, argv, "dthr", 2);
    c->boost = argu(argc, argv, "boost", 4);
    c->mode = argu(argc, argv, "mode", 1);
    c->epochs = argu(argc, argv, "epochs", 30);
    c->T = argu(argc, argv, "T", 10);
    c->clampsteps = argu(argc, argv, "clampsteps", 5);
    c->stim = argu(argc, argv, "stim", 12);
    c->quiet = argu(argc, argv, "quiet", 0);
    c->seed = argu(argc, argv, "seed", 0);
    c->split = argu(argc, argv, "split", 0);
    c->unionchar = argu(argc, argv, "unionchar", 0);
    c->biasup = argu(argc, argv, "biasup", 0);
    c->biasdown = argu(argc, argv, "biasdown", 1);
    c->biasshift = argu(argc, argv, "biasshift", 4);
    c->biasdownevery = argu(argc, argv, "biasdownevery", 1);
    c->inbr = argu(argc, argv, "inbr", 8);
    c->stimcols = argu(argc, argv, "stimcols", 128);
    c->incols = argu(argc, argv, "incols", 512);
    c->ncol = argu(argc, argv, "ncol", 4096);
    c->diag = argu(argc, argv, "diag", 0);
    c->profile = argu(argc, argv, "profile", 0);
    c->topkpar = argu(argc, argv, "topkpar", 1);
    c->vtopk = argu(argc, argv, "vtopk", 0);
    c->contrast = argu(argc, argv, "contrast", 0);
    c->bandlo = argu(argc, argv, "bandlo", 5);
    c->bandhi = argu(argc, argv, "bandhi", 11);
    c->margingate = argu(argc, argv, "margingate", 1);
    c->hetero = argu(argc, argv, "hetero", 0);
    c->vbs = argu(argc, argv, "vbs", 0);
    c->ltdrep = argu(argc, argv, "ltdrep", 1);
    c->ltdevery = argu(argc, argv, "ltdevery", 1);
    c->ktest = argu(argc, argv, "ktest", 0);
    c->facil = argu(argc, argv, "facil", 0);
    c->deep = argu(argc, argv, "deep", 0);
    c->wave = argu(argc, argv, "wave", 0);
    c->wavet = argu(argc, argv, "wavet", 35);
    c->trainclamp = argu(argc, argv, "trainclamp", c->contrast ? c->clampsteps : c->T);
    if (c->biasup == 1u) {  // mm:732-735
        double pr = (double)c->target * c->k / (double)(c->ncol - c->incols) / NPC;
        c->biasup = (uint32_t)std::max(2.0, (1.0 - pr) / pr * c->biasdown);
    }
    c->max_t = argu(argc, argv, "maxt", 64);
}

extern "C" int ora_config_check(const ora_config* c, char* msg, size_t len) {
    auto bad = [&](const char* m) { if (msg && len) snprintf(msg, len, "%s", m); return 1; };
    if (c->ncol == 0 || c->ncol % COLS_PER_GRP) return bad("ncol must be a positive multiple of 8");
    if (c->incols == 0 || c->incols % COLS_PER_GRP) return bad("incols must be a positive multiple of 8");
    if (c->incols >= c->ncol) return bad("incols must be < ncol");
    if (c->stimcols > c->incols) return bad("stimcols must be <= incols");
    if (c->inbr > NSLOT) return bad("inbr must be <= 32");
    if (c->biasshift >= 32) return bad("biasshift must be < 32");
    if (c->ltdevery == 0 && c->contrast) return bad("ltdevery must be >= 1");
    if (c->max_t == 0) return bad("max_t must be >= 1");
    if (c->stim == 0) return bad("stim must be >= 1");
    return 0;
}

// ------------------------------------------------------------------ init
// k_init_w (mm:86-104) and k_init_conn (mm:109-117); stimulus columns mm:793-799
extern "C" ora_net* ora_net_create(const ora_config* c, int nthreads) {
    if (ora_config_check(c, nullptr, 0)) return nullptr;
    ora_net* n = new ora_net();
    n->cfg = *c;
    n->nthreads = nthreads;
    n->NCOL = c->ncol; n->N = c->ncol * NPC; n->NGRP = c->ncol / COLS_PER_GRP; n->IN_COLS = c->incols;
    n->MAXT = c->max_t;
    n->NSTATS = std::max<uint32_t>(1024u, 64u + 8u * n->MAXT);
    n->NDLOG = std::max<uint32_t>(256u, 4u * n->MAXT);
    n->PLANE_U4 = (size_t)n->NCOL * NSLOT * SLOT_U4 * 32;
    for (int p = 0; p < 4; ++p) n->W[p].assign(n->PLANE_U4 * 4, 0u);
    n->conn.assign((size_t)n->NCOL * NSLOT, 0u);
    n->spkA.assign(n->NCOL, 0u); n->spkB.assign(n->NCOL, 0u); n->drv.assign(n->NCOL, 0u);
    n->bbits.assign(n->N, 0u); n->acc.assign(n->N, 0u); n->bias.assign(n->N, 0u);
    n->marg.assign(n->NCOL, 0u); n->refBall.assign(n->NCOL, 0u); n->driveLog.assign(n->NDLOG, 0u);
    n->mult.assign(n->N, 0u); n->stim.assign(n->NCOL, 0u); n->stat.assign(n->NSTATS, 0u);
    n->hist.assign((size_t)n->MAXT * n->NCOL, 0u); n->theta.assign(16, 0u);
    n->cntProp.assign(n->NCOL, 0); n->cntLearn.assign(n->NCOL, 0);

    const uint32_t seed = c->seed, nAnd = c->nand;
    uint32_t* W0 = n->W[0].data(); uint32_t* W1 = n->W[1].data();
    uint32_t* W2 = n->W[2].data(); uint32_t* W3 = n->W[3].data();
    for (size_t gg64 = 0; gg64 < n->PLANE_U4; ++gg64) {
        uint32_t gg = (uint32_t)gg64;
        uint32_t g = gg ^ (seed * 0x9E3779B9u);
        uint32_t m[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
        for (uint32_t i = 0; i < nAnd; ++i)
            for (uint32_t q = 0; q < 4; ++q) m[q] &= hash32(g * 32u + i * 4u + q);
        const uint32_t cc[4] = {hash32(g * 8u + 11), hash32(g * 8u + 13), hash32(g * 8u + 17),
                                hash32(g * 8u + 19)};
        for (uint32_t q = 0; q < 4; ++q) {
            size_t w = gg64 * 4 + q;
            W3[w] = m[q];
            W2[w] = (~m[q]) | cc[q];
            W1[w] = cc[q];
            W0[w] = ~cc[q];
        }
    }
    const uint32_t nGrp = n->NGRP, nInGrp = n->IN_COLS / COLS_PER_GRP, inBr = c->inbr;
    for (uint32_t g = 0; g < n->NCOL * NSLOT; ++g) {
        uint32_t s = g % NSLOT;
        uint32_t h = hash32(g * 2246822519u + 7u + seed * 2654435761u);
        n->conn[g] = (s < inBr) ? (h % nInGrp) : (nInGrp + h % (nGrp - nInGrp));
    }
    std::mt19937 rng(7u + seed * 104729u);
    n->stimCols.resize(c->stim);
    std::vector<uint32_t> all(n->IN_COLS);
    for (uint32_t p = 0; p < c->stim; ++p) {
        for (uint32_t i = 0; i < n->IN_COLS; ++i) all[i] = i;
        libcxx_shuffle(all.data(), n->IN_COLS, rng);
        n->stimCols[p].assign(all.begin(), all.begin() + c->stimcols);
    }
    // P fields fixed at setup, mm:810-812
    Params& P = n->P;
    memset(&P, 0, sizeof(P));
    P.k = c->k; P.dthr = c->dthr; P.boost = c->boost; P.mode = c->mode; P.inCols = c->incols;
    P.targetCols = c->target; P.biasUp = c->biasup; P.biasDown = c->biasdown;
    P.biasShift = c->biasshift; P.bandLo = c->bandlo; P.bandHi = c->contrast ? c->bandhi : 0u;
    P.marginGate = c->margingate; P.hetero = c->hetero; P.ltdRep = c->ltdrep;
    P.biasDownEvery = c->biasdownevery;
    n->K = c->k; n->kTest = c->ktest; n->facil = c->facil; n->contrast = c->contrast;
    n->ltdEvery = c->ltdevery; n->topkPar = c->topkpar;
    return n;
}

extern "C" void ora_net_destroy(ora_net* n) { delete n; }
extern "C" void ora_net_set_threads(ora_net* n, int t) { n->nthreads = t; }

// ------------------------------------------------------------------ kernels
// k_drive, mm:150-190
static void k_drive(ora_net* n, const uint32_t* spkIn) {
    const Params P = n->P;
    const uint32_t* W3 = n->W[3].data();
    par_cols(n->nthreads, 0, n->NCOL, [&](uint32_t c0, uint32_t c1, int) {
        for (uint32_t col = c0; col < c1; ++col) {
            uint32_t acc[32] = {0}, bb[32] = {0};
            const uint32_t base = col * NSLOT;
            for (uint32_t s = 0; s < NSLOT; ++s) {
                const uint32_t g = n->conn[base + s];
                const uint32_t* sp = spkIn + (size_t)g * 8;
                uint32_t any = sp[0] | sp[1] | sp[2] | sp[3] | sp[4] | sp[5] | sp[6] | sp[7];
                if (any == 0u) continue;
                n->cntProp[col]++;
                for (uint32_t lane = 0; lane < 32; ++lane) {
                    const size_t wb = ((size_t)(base + s) * SLOT_U4 * 32u + lane) * 4;
                    const uint32_t* w0 = W3 + wb;           // uint4 at wb
                    const uint32_t* w1 = W3 + wb + 32 * 4;  // uint4 at wb + 32
                    uint32_t c = popc(w0[0] & sp[0]) + popc(w0[1] & sp[1]) + popc(w0[2] & sp[2]) +
                                 popc(w0[3] & sp[3]) + popc(w1[0] & sp[4]) + popc(w1[1] & sp[5]) +
                                 popc(w1[2] & sp[6]) + popc(w1[3] & sp[7]);
                    bool nmda = (c >= P.dthr);
                    if (nmda) bb[lane] |= (1u << s);
                    if (P.mode == 0u) acc[lane] += c;
                    else if (P.mode == 1u) acc[lane] += nmda ? (c + P.boost) : c;
                    else acc[lane] += nmda ? 1u : 0u;
                }
            }
            uint32_t drv = 0;
            for (uint32_t lane = 0; lane < 32; ++lane) {
                const uint32_t gid = col * 32 + lane;
                uint32_t a = acc[lane];
                if (P.facil != 0u && ((spkIn[col] >> lane) & 1u) != 0u) a += P.facil;
                uint32_t bi = (P.useBias != 0u) ? (n->bias[gid] >> P.biasShift) : 0u;
                a = (a > bi) ? (a - bi) : 0u;
                n->acc[gid] = a;
                n->bbits[gid] = bb[lane];
                drv += a;
            }
            n->drv[col] = drv;
        }
    });
}

// k_emit, mm:192-236
static void k_emit(ora_net* n, uint32_t* spkOut) {
    const Params P = n->P;
    const uint32_t th = std::max(n->theta[0], 1u);
    int nt = n->nthreads > 1 ? n->nthreads : 1;
    std::vector<std::vector<uint32_t>> lstat(nt, std::vector<uint32_t>(n->NSTATS, 0u));
    std::vector<std::vector<uint32_t>> ldl(nt, std::vector<uint32_t>(n->NDLOG, 0u));
    par_cols(n->nthreads, 0, n->NCOL, [&](uint32_t c0, uint32_t c1, int tid) {
        uint32_t* st = lstat[tid].data();
        uint32_t* dl = ldl[tid].data();
        for (uint32_t col = c0; col < c1; ++col) {
            uint32_t v[32];
            for (uint32_t l = 0; l < 32; ++l) v[l] = n->acc[col * 32 + l];
            uint32_t vk = 0, vk1 = 0;
            uint32_t b = kwta_margin(v, P.k, vk, vk1);
            if (n->drv[col] < th) b = 0u;
            if (P.clamp != 0u && col < P.inCols) b = n->stim[col];
            for (uint32_t l = 0; l < 32; ++l) {
                const uint32_t gid = col * 32 + l;
                if (((b >> l) & 1u) != 0u && col >= P.inCols) st[2] += popc(n->bbits[gid]);
                if (col >= P.inCols && ((n->refBall[col] >> l) & 1u) != 0u) {
                    uint32_t o = (n->mult[gid] > 1u) ? 0u : 2u;
                    dl[P.step * 4u + o + 0u] += n->acc[gid];
                    dl[P.step * 4u + o + 1u] += 1u;
                }
            }
            spkOut[col] = b;
            n->marg[col] = (vk > vk1) ? 1u : 0u;
            if (P.histOn != 0u) n->hist[(size_t)P.step * n->NCOL + col] = b;
            if (b != 0u && col >= P.inCols) {
                uint32_t sb = 64u + P.step * 8u;
                st[0] += 1u; st[1] += popc(b);
                st[sb + 0] += vk - vk1; st[sb + 1] += vk; st[sb + 2] += 1u;
                if (vk == vk1) st[sb + 3] += 1u;
                if (vk == 0u) st[sb + 4] += 1u;
            }
        }
    });
    for (int t = 0; t < nt; ++t) {
        for (uint32_t i = 0; i < n->NSTATS; ++i) n->stat[i] += lstat[t][i];
        for (uint32_t i = 0; i < n->NDLOG; ++i) n->driveLog[i] += ldl[t][i];
    }
}

// k_bias, mm:240-253
static void k_bias(ora_net* n, const uint32_t* spk) {
    const Params P = n->P;
    for (uint32_t gid = P.inCols * 32u; gid < n->N; ++gid) {
        uint32_t col = gid >> 5, lane = gid & 31u;
        bool fired = ((spk[col] >> lane) & 1u) != 0u;
        int dec = ((P.step % std::max(P.biasDownEvery, 1u)) == 0u) ? (int)P.biasDown : 0;
        int b = (int)n->bias[gid] + (fired ? (int)P.biasUp : -dec);
        n->bias[gid] = (uint32_t)std::max(0, b);
    }
}

// k_learn, mm:436-495
static void k_learn(ora_net* n, const uint32_t* spkIn, const uint32_t* spkOut) {
    const Params P = n->P;
    if (P.phase == 3u) return;
    int nt = n->nthreads > 1 ? n->nthreads : 1;
    std::vector<std::array<uint32_t, 4>> lst(nt, std::array<uint32_t, 4>{0, 0, 0, 0});
    uint32_t* Wp[4] = {n->W[0].data(), n->W[1].data(), n->W[2].data(), n->W[3].data()};
    par_cols(n->nthreads, 0, n->NCOL, [&](uint32_t c0, uint32_t c1, int tid) {
        auto& st = lst[tid];
        for (uint32_t col = c0; col < c1; ++col) {
            uint32_t b = spkOut[col];
            if (b == 0u) continue;
            if (P.marginGate != 0u && n->marg[col] == 0u) continue;
            uint32_t nLTP[32] = {0}, nLTD[32] = {0};
            const uint32_t base = col * NSLOT;
            for (uint32_t s = 0; s < NSLOT; ++s) {
                uint32_t vote = 0;
                for (uint32_t l = 0; l < 32; ++l) {
                    bool post = ((b >> l) & 1u) != 0u;
                    bool want = post && (((n->bbits[col * 32 + l] >> s) & 1u) != 0u);
                    if (want) vote |= 1u << l;
                }
                if (vote == 0u) continue;
                n->cntLearn[col]++;
                const uint32_t g = n->conn[base + s];
                for (uint32_t l = 0; l < 32; ++l) {
                    const bool want = ((vote >> l) & 1u) != 0u;
                    for (uint32_t j = 0; j < SLOT_U4; ++j) {
                        const uint32_t* sp = spkIn + (size_t)g * 8 + j * 4;
                        const size_t idx = ((size_t)(base + s) * SLOT_U4 * 32u + j * 32u + l) * 4;
                        for (uint32_t q = 0; q < 4; ++q) {
                            uint32_t a0 = Wp[0][idx + q], a1 = Wp[1][idx + q], a2 = Wp[2][idx + q],
                                     a3 = Wp[3][idx + q];
                            uint32_t co = want ? sp[q] : 0u;
                            uint32_t nco = popc(co);
                            if (P.phase == 1u) { bs_inc(a0, a1, a2, a3, co); nLTP[l] += nco; }
                            else if (P.phase == 2u) {
                                for (uint32_t rr = 0; rr < P.ltdRep; ++rr) bs_dec(a0, a1, a2, a3, co);
                                nLTD[l] += nco * P.ltdRep;
                            } else {
                                bs_inc(a0, a1, a2, a3, co);
                                bs_dec(a0, a1, a2, a3, (want && P.depress != 0u) ? ~sp[q] : 0u);
                            }
                            if (P.hetero != 0u && P.phase == 1u) bs_dec(a0, a1, a2, a3, want ? ~sp[q] : 0u);
                            if (P.bandHi != 0u) bs_clip(a0, a1, a2, a3, P.bandLo, P.bandHi);
                            Wp[0][idx + q] = a0; Wp[1][idx + q] = a1; Wp[2][idx + q] = a2; Wp[3][idx + q] = a3;
                        }
                    }
                }
            }
            for (uint32_t l = 0; l < 32; ++l) {
                if (nLTP[l]) st[0] += nLTP[l];
                if (nLTD[l]) st[1] += nLTD[l];
                if (nLTP[l]) st[2] += 1u;
                if (nLTD[l]) st[3] += 1u;
            }
        }
    });
    for (int t = 0; t < nt; ++t)
        for (int i = 0; i < 4; ++i) n->stat[8 + i] += lst[t][i];
}

// ------------------------------------------------------------------ present
// loadStim, mm:800-804
static void load_stim(ora_net* n, uint32_t p, double keep, uint32_t seed) {
    std::fill(n->stim.begin(), n->stim.end(), 0u);
    std::mt19937 r(seed);
    for (uint32_t c : n->stimCols[p])
        if (uniform01(r) <= keep) n->stim[c] = 0xFFFFFFFFu;
}

// present(), mm:815-888
extern "C" uint32_t ora_present(ora_net* n, uint32_t p, double keep, uint32_t T_, uint32_t clampSteps,
                                uint32_t learn, uint32_t seed, uint32_t useBias,
                                const uint32_t* stim_override) {
    if (T_ > n->MAXT) return 0;
    if (!stim_override && p >= n->stimCols.size()) return 0;
    Params& P = n->P;
    P.k = learn ? n->K : (n->kTest != 0u ? n->kTest : n->K);
    P.facil = learn ? 0u : n->facil;
    P.useBias = (useBias == 2u) ? (learn ? 1u : 0u) : useBias;
    if (stim_override) std::copy(stim_override, stim_override + n->NCOL, n->stim.begin());
    else load_stim(n, p, keep, seed);
    std::fill(n->stat.begin(), n->stat.end(), 0u);
    P.histOn = 1u;
    std::fill(n->spkA.begin(), n->spkA.end(), 0u);
    std::fill(n->spkB.begin(), n->spkB.end(), 0u);
    for (uint32_t t = 0; t < T_; ++t) {
        uint32_t* in = (t & 1) ? n->spkB.data() : n->spkA.data();
        uint32_t* out = (t & 1) ? n->spkA.data() : n->spkB.data();
        P.step = t; P.learn = learn ? 1u : 0u; P.freeze = learn ? 0u : 1u;
        P.clamp = (t < clampSteps) ? 1u : 0u;
        P.depress = (learn && (t % 2 == 1)) ? 1u : 0u;
        P.phase = (!learn || !n->contrast)
                      ? 0u
                      : ((t < clampSteps) ? 1u : (((t - clampSteps) % n->ltdEvery == 0u) ? 2u : 3u));
        k_drive(n, in);
        uint32_t th[2];
        topk(n->drv.data(), n->NCOL, P.inCols, P.targetCols, n->topkPar != 0, th);
        n->theta[0] = th[0]; n->theta[1] = th[1];
        k_emit(n, out);
        if (learn) {
            k_bias(n, out);
            k_learn(n, in, out);
        }
    }
    return T_;
}

extern "C" uint32_t* ora_buffer(ora_net* n, int which, size_t* nw) {
    std::vector<uint32_t>* v = nullptr;
    switch (which) {
        case ORA_W0: v = &n->W[0]; break;
        case ORA_W1: v = &n->W[1]; break;
        case ORA_W2: v = &n->W[2]; break;
        case ORA_W3: v = &n->W[3]; break;
        case ORA_CONN: v = &n->conn; break;
        case ORA_SPK_A: v = &n->spkA; break;
        case ORA_SPK_B: v = &n->spkB; break;
        case ORA_DRIVE: v = &n->drv; break;
        case ORA_BBITS: v = &n->bbits; break;
        case ORA_ACC: v = &n->acc; break;
        case ORA_BIAS: v = &n->bias; break;
        case ORA_MARG: v = &n->marg; break;
        case ORA_REFBALL: v = &n->refBall; break;
        case ORA_DRIVELOG: v = &n->driveLog; break;
        case ORA_MULT: v = &n->mult; break;
        case ORA_STIM: v = &n->stim; break;
        case ORA_STATS: v = &n->stat; break;
        case ORA_HIST: v = &n->hist; break;
        case ORA_THETA: v = &n->theta; break;
        default: if (nw) *nw = 0; return nullptr;
    }
    if (nw) *nw = v->size();
    return v->data();
}

extern "C" const uint32_t* ora_final_spikes(ora_net* n, uint32_t T) {
    return (T & 1) ? n->spkB.data() : n->spkA.data();
}

extern "C" void ora_stim_cols(ora_net* n, uint32_t p, uint32_t* out) {
    std::copy(n->stimCols[p].begin(), n->stimCols[p].end(), out);
}

// ------------------------------------------------------------------ readouts
// k_within, mm:524-562
extern "C" void ora_within(ora_net* n, const uint32_t* ball, const uint32_t* postMask, uint32_t out[16]) {
    const uint32_t inBr = n->cfg.inbr, inCols = n->IN_COLS;
    const uint32_t* W3 = n->W[3].data();
    memset(out, 0, 16 * sizeof(uint32_t));  // bWCnt is zeroed before each call (mm:1701)
    for (uint32_t gid = 0; gid < n->N; ++gid) {
        uint32_t col = gid >> 5, lane = gid & 31u;
        if (col < inCols) continue;
        if (postMask) { if (postMask[gid] == 0u) continue; }
        else if (((ball[col] >> lane) & 1u) == 0u) continue;
        uint32_t base = col * NSLOT, fIn = 0, pIn = 0, fOut = 0, pOut = 0;
        for (uint32_t s = inBr; s < NSLOT; ++s) {
            uint32_t g = n->conn[base + s];
            for (uint32_t j = 0; j < SLOT_U4; ++j) {
                const uint32_t* ap = ball + (size_t)g * 8 + j * 4;
                const uint32_t* w = W3 + ((size_t)(base + s) * SLOT_U4 * 32u + j * 32u + lane) * 4;
                for (int q = 0; q < 4; ++q) {
                    fIn += popc(w[q] & ap[q]); pIn += popc(ap[q]);
                    fOut += popc(w[q] & ~ap[q]); pOut += popc(~ap[q]);
                }
            }
        }
        out[0] += fIn; out[1] += pIn; out[2] += fOut; out[3] += pOut;
        uint32_t b = (n->mult[gid] > 1u) ? 4u : 6u;
        out[b + 0] += fIn; out[b + 1] += pIn;
        out[(n->mult[gid] > 1u) ? 8u : 9u] += 1u;
    }
}

// k_wstat, mm:641-646
extern "C" uint32_t ora_wstat(ora_net* n) {
    uint32_t s = 0;
    for (uint32_t w : n->W[3]) s += popc(w);
    return s;
}

// k_indeg, mm:566-581
extern "C" void ora_indeg(ora_net* n, uint32_t* out) {
    for (uint32_t gid = 0; gid < n->N; ++gid) {
        uint32_t col = gid >> 5, lane = gid & 31u, d = 0;
        for (uint32_t s = n->cfg.inbr; s < NSLOT; ++s)
            for (uint32_t j = 0; j < SLOT_U4; ++j) {
                const uint32_t* w = n->W[3].data() + ((size_t)(col * NSLOT + s) * 64u + j * 32u + lane) * 4;
                d += popc(w[0]) + popc(w[1]) + popc(w[2]) + popc(w[3]);
            }
        out[gid] = d;
    }
}

// k_untouched, mm:501-519
extern "C" void ora_untouched(ora_net* n, uint32_t* out) {
    for (uint32_t gid = 0; gid < n->N; ++gid) {
        uint32_t col = gid >> 5, lane = gid & 31u, u = 0;
        for (uint32_t s = n->cfg.inbr; s < NSLOT; ++s)
            for (uint32_t j = 0; j < SLOT_U4; ++j) {
                size_t idx = ((size_t)(col * NSLOT + s) * 64u + j * 32u + lane) * 4;
                for (int q = 0; q < 4; ++q) {
                    uint32_t a0 = n->W[0][idx + q], a1 = n->W[1][idx + q], a2 = n->W[2][idx + q],
                             a3 = n->W[3][idx + q];
                    uint32_t v5 = (~a3) & a2 & (~a1) & a0, v6 = (~a3) & a2 & a1 & (~a0);
                    u += popc(v5 | v6);
                }
            }
        out[gid] = u;
    }
}

// k_wdiag, mm:591-638
extern "C" void ora_wdiag(ora_net* n, const uint32_t* memb, const uint32_t* asmSpk, uint32_t nstim,
                          uint32_t sel, uint32_t hist[48]) {
    const uint32_t inBr = n->cfg.inbr, ncol4 = n->NCOL / 4;
    memset(hist, 0, 48 * sizeof(uint32_t));  // zeroed before each call (mm:1127)
    for (uint32_t gid = 0; gid < n->N; ++gid) {
        uint32_t mypost = memb[gid];
        if (mypost == 0u) continue;
        uint32_t col = gid >> 5, lane = gid & 31u, base = col * NSLOT;
        uint32_t acc[48] = {0};
        for (uint32_t s = 0; s < NSLOT; ++s) {
            bool isAff = (s < inBr);
            if (sel == 0u && isAff) continue;
            if (sel == 1u && !isAff) continue;
            uint32_t g = n->conn[base + s];
            for (uint32_t j = 0; j < SLOT_U4; ++j) {
                size_t idx = ((size_t)(base + s) * 64u + j * 32u + lane) * 4;
                for (int q = 0; q < 4; ++q) {
                    uint32_t a0 = n->W[0][idx + q], a1 = n->W[1][idx + q], a2 = n->W[2][idx + q],
                             a3 = n->W[3][idx + q];
                    uint32_t within = 0, anyp = 0;
                    for (uint32_t st = 0; st < nstim; ++st) {
                        uint32_t ap = asmSpk[((size_t)st * ncol4 + 2u * g + j) * 4 + q];
                        anyp |= ap;
                        if (((mypost >> st) & 1u) != 0u) within |= ap;
                    }
                    uint32_t between = anyp & ~within;
                    for (uint32_t v = 0; v < 16u; ++v) {
                        uint32_t eq = ((v & 1u) ? a0 : ~a0) & ((v & 2u) ? a1 : ~a1) &
                                      ((v & 4u) ? a2 : ~a2) & ((v & 8u) ? a3 : ~a3);
                        acc[v] += popc(eq & within);
                        acc[16 + v] += popc(eq & between);
                        acc[32 + v] += popc(eq);
                    }
                }
            }
        }
        for (int i = 0; i < 48; ++i) hist[i] += acc[i];
    }
}

extern "C" void ora_bs_apply(uint32_t op, uint32_t lo, uint32_t hi, uint32_t* a0, uint32_t* a1,
                             uint32_t* a2, uint32_t* a3, const uint32_t* m, size_t nw) {
    for (size_t i = 0; i < nw; ++i) {
        if (op == 0) bs_inc(a0[i], a1[i], a2[i], a3[i], m[i]);
        else if (op == 1) bs_dec(a0[i], a1[i], a2[i], a3[i], m[i]);
        else bs_clip(a0[i], a1[i], a2[i], a3[i], lo, hi);
    }
}

extern "C" void ora_topk(const uint32_t* drive, uint32_t ncol, uint32_t incols, uint32_t target, int par,
                         uint32_t theta[2]) {
    topk(drive, ncol, incols, target, par != 0, theta);
}

extern "C" uint32_t ora_kwta(const uint32_t v[32], uint32_t k, uint32_t* vk, uint32_t* vk1) {
    uint32_t a, b;
    uint32_t r = kwta_margin(v, k, a, b);
    if (vk) *vk = a;
    if (vk1) *vk1 = b;
    return r;
}

extern "C" uint32_t ora_mt_nth(uint32_t seed, uint32_t nth) {
    std::mt19937 g(seed);
    uint32_t v = 0;
    for (uint32_t i = 0; i < nth; ++i) v = (uint32_t)g();
    return v;
}

extern "C" void ora_libcxx_shuffle(uint32_t* a, uint32_t n, uint32_t mt_seed) {
    std::mt19937 g(mt_seed);
    libcxx_shuffle(a, n, g);
}

extern "C" double ora_uniform01(uint32_t seed, uint32_t idx) {
    std::mt19937 g(seed);
    double u = 0;
    for (uint32_t i = 0; i <= idx; ++i) u = uniform01(g);
    return u;
}

extern "C" void ora_work_counters(ora_net* n, uint64_t out[2], int reset) {
    out[0] = 0; out[1] = 0;
    for (uint64_t x : n->cntProp) out[0] += x;
    for (uint64_t x : n->cntLearn) out[1] += x;
    if (reset) {
        std::fill(n->cntProp.begin(), n->cntProp.end(), 0);
        std::fill(n->cntLearn.begin(), n->cntLearn.end(), 0);
    }
}
