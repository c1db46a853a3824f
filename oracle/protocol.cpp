// AXON CPU oracle protocol -- TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Restates the reference main() after the network is built (mm:996-1911):
// every presentation in the reference's order and every report block it
// prints, on the oracle's kernels.  Left out: the Metal timestamp profile
// (profile=1, mm:1162-1340), the vbs / vtopk self-checks (mm:903-994; their
// own tests cover them), and the lines that carry a measured time (the
// ms/timestep line mm:1154 and the quiet one-liner mm:1907).  Blocks:
//   initial density + settle      mm:996-1005
//   epochs                        mm:1008-1024
//   firing steps (extra present)  mm:1028-1072
//   references + cross-talk       mm:1080-1105
//   STEP 0 diagnostic (diag)      mm:1107-1151
//   STEP A1                       mm:1342-1409
//   STEP A3 + A4                  mm:1411-1511
//   A4b + A4c                     mm:1513-1599
//   UNIONCHAR (unionchar)         mm:1601-1652
//   SHARED vs EXCLUSIVE (split)   mm:1654-1692
//   scaffold density (deep)       mm:1694-1730
//   multiplicity (deep)           mm:1732-1801
//   PARITY + WAVE (wave)          mm:1803-1858
//   free-run diagnostic           mm:1860-1887
//   final completion table        mm:1889-1906
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "oracle.h"

namespace {

struct Out {
    std::string s;
    void f(const char* fmt, ...) __attribute__((format(printf, 2, 3))) {
        char b[8192];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(b, sizeof b, fmt, ap);
        va_end(ap);
        s += b;
    }
};

using Vec = std::vector<uint32_t>;

// overlap, mm:896-901
double overlap(const Vec& a, const Vec& b) {
    if (a.empty() || b.empty()) return 0.0;
    Vec in;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(in));
    return (double)in.size() / std::sqrt((double)a.size() * b.size());
}
// the Jaccard lambdas of mm:1382-1387 and mm:1517-1521 (0 when either set is empty)
double jacc(const Vec& a, const Vec& b) {
    if (a.empty() || b.empty()) return 0.0;
    Vec i, u;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(i));
    std::set_union(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(u));
    return (double)i.size() / u.size();
}
// jac2 of mm:1776-1780 (0 only when the union is empty)
double jacc_union(const Vec& a, const Vec& b) {
    Vec i, u;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(i));
    std::set_union(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(u));
    return u.empty() ? 0.0 : (double)i.size() / u.size();
}
Vec cols_of(const uint32_t* s, uint32_t inCols, uint32_t ncol) {
    Vec v;
    for (uint32_t c = inCols; c < ncol; ++c)
        if (s[c]) v.push_back(c);
    return v;
}
Vec neuron_set(const uint32_t* bal, uint32_t inCols, uint32_t ncol) {
    Vec r;
    for (uint32_t c = inCols; c < ncol; ++c) {
        uint32_t b = bal[c];
        while (b) { uint32_t l = __builtin_ctz(b); b &= b - 1; r.push_back(c * 32 + l); }
    }
    return r;
}
// Pearson correlation, the pear lambda of mm:1793-1797
double pearson(const std::vector<double>& x, const std::vector<double>& y) {
    double mx = 0, my = 0;
    for (size_t i = 0; i < x.size(); ++i) { mx += x[i]; my += y[i]; }
    mx /= x.size(); my /= y.size();
    double sxy = 0, sxx = 0, syy = 0;
    for (size_t i = 0; i < x.size(); ++i) {
        double dx = x[i] - mx, dy = y[i] - my;
        sxy += dx * dy; sxx += dx * dx; syy += dy * dy;
    }
    return (sxx > 0 && syy > 0) ? sxy / std::sqrt(sxx * syy) : 0.0;
}

}  // namespace

extern "C" int ora_run_protocol(const ora_config* cin, int nthreads, ora_pinned* pin, char* buf,
                                size_t buf_len) {
    ora_config c = *cin;
    Out o;
    ora_net* n = ora_net_create(&c, nthreads);
    if (!n) return 1;
    const uint32_t NCOL = c.ncol, N = NCOL * 32, IN = c.incols, NSTIM = c.stim, T = c.T, CLAMP = c.clampsteps;
    const uint32_t trainClamp = c.trainclamp, EPOCHS = c.epochs, inBranches = c.inbr, biasUp = c.biasup;
    const uint32_t MAXT = c.max_t ? c.max_t : 64u;
    const bool quiet = c.quiet != 0;
    memset(pin, 0, sizeof(*pin));
    size_t nw = 0;
    uint32_t* stats = ora_buffer(n, ORA_STATS, &nw);
    uint32_t* hist = ora_buffer(n, ORA_HIST, &nw);
    uint32_t* theta = ora_buffer(n, ORA_THETA, &nw);
    uint32_t* bias = ora_buffer(n, ORA_BIAS, &nw);
    uint32_t* refBall = ora_buffer(n, ORA_REFBALL, &nw);
    uint32_t* multBuf = ora_buffer(n, ORA_MULT, &nw);
    uint32_t* dl = ora_buffer(n, ORA_DRIVELOG, &nw);
    const size_t ndl = nw;
    const double NSYN = (double)N * 8192.0;
    auto wstat = [&]() { return 100.0 * ora_wstat(n) / NSYN; };  // mm:784-791
    auto final_spikes = [&](uint32_t TT) { return ora_final_spikes(n, TT); };
    std::vector<Vec> stimCols(NSTIM, Vec(c.stimcols));
    for (uint32_t p = 0; p < NSTIM; ++p) ora_stim_cols(n, p, stimCols[p].data());

    // initial density and settling, mm:996-1005
    const double d0 = wstat();
    if (!quiet) o.f("initial functional synapse density: %.3f%%  (fan-in %.0f functional)\n\n", d0, d0 / 100.0 * 8192);
    for (uint32_t i = 0; i < 10; ++i) ora_present(n, i % NSTIM, 1.0, T, T, 0, 999, 2, nullptr);
    if (!quiet)
        o.f("settled: theta=%u  activeCols/step=%.0f  activeNeurons/step=%.0f\n\n", theta[0], (double)stats[0] / T,
            (double)stats[1] / T);

    // epochs, mm:1008-1024
    for (uint32_t ep = 0; ep < EPOCHS; ++ep) {
        for (uint32_t p = 0; p < NSTIM; ++p) ora_present(n, p, 1.0, T, trainClamp, 1, 1000 + p, 2, nullptr);
        if (!quiet && (ep % 5 == 4 || ep == 0)) {
            double ltp = stats[8], ltd = stats[9];
            uint32_t tp = stats[10], td = stats[11];
            double wd = wstat();
            o.f("  ep %3u  wdens %7.3f%%  LTP %10.0f  LTD %10.0f  LTD/LTP %6.3f  learners %u/%u\n", ep + 1, wd,
                ltp, ltd, ltp > 0 ? ltd / ltp : 0.0, tp, td);
        }
    }
    if (EPOCHS == 0 || NSTIM == 0) ora_present(n, 0, 1.0, T, T, 0, 777, 2, nullptr);  // timed, non-learning

    // firing steps per neuron over one extra learning presentation, mm:1028-1072
    {
        ora_present(n, 0, 1.0, T, trainClamp, 1, 31337, 2, nullptr);
        std::vector<uint16_t> cnt(N, 0), cntC(N, 0), cntF(N, 0);
        for (uint32_t t = 0; t < T; ++t)
            for (uint32_t col = IN; col < NCOL; ++col) {
                uint32_t b = hist[(size_t)t * NCOL + col];
                while (b) {
                    uint32_t l = __builtin_ctz(b); b &= b - 1;
                    cnt[col * 32 + l]++;
                    if (t < trainClamp) cntC[col * 32 + l]++; else cntF[col * 32 + l]++;
                }
            }
        const uint32_t* last = hist + (size_t)(T - 1) * NCOL;
        std::vector<uint8_t> member(N, 0);
        for (uint32_t col = IN; col < NCOL; ++col) {
            uint32_t b = last[col];
            while (b) { uint32_t l = __builtin_ctz(b); b &= b - 1; member[col * 32 + l] = 1; }
        }
        double nn = 0, sum = 0, sumC = 0, sumF = 0, nm = 0, sm = 0, smC = 0, smF = 0;
        std::vector<uint16_t> v, vm;
        for (uint32_t i = IN * 32; i < N; ++i) {
            if (cnt[i]) { nn++; sum += cnt[i]; sumC += cntC[i]; sumF += cntF[i]; v.push_back(cnt[i]); }
            if (member[i]) { nm++; sm += cnt[i]; smC += cntC[i]; smF += cntF[i]; vm.push_back(cnt[i]); }
        }
        std::sort(v.begin(), v.end());
        std::sort(vm.begin(), vm.end());
        o.f("\n================= FIRING STEPS PER NEURON (training) =================\n");
        o.f("  T=%u  trainClamp=%u (clamped) + %u (free)\n", T, trainClamp, T - trainClamp);
        o.f("  neurons firing at least once: %.0f\n", nn);
        o.f("  mean firing steps  : %.2f   (clamped %.2f + free %.2f)\n", nn ? sum / nn : 0.0, nn ? sumC / nn : 0.0,
            nn ? sumF / nn : 0.0);
        if (!v.empty())
            o.f("  p50 %u | p90 %u | max %u   | as fraction of T: %.3f\n", v[v.size() / 2],
                v[(size_t)(0.9 * (v.size() - 1))], v.back(), nn ? sum / nn / T : 0.0);
        o.f("  --- restricted to FINAL-STATE ASSEMBLY MEMBERS (the allocated ones) ---\n");
        o.f("  members: %.0f | mean firing steps %.2f (clamped %.2f + free %.2f) | fraction of T %.3f\n", nm,
            nm ? sm / nm : 0.0, nm ? smC / nm : 0.0, nm ? smF / nm : 0.0, nm ? sm / nm / T : 0.0);
        if (!vm.empty())
            o.f("  p50 %u | p90 %u | max %u\n", vm[vm.size() / 2], vm[(size_t)(0.9 * (vm.size() - 1))], vm.back());
        o.f("  => bias gain per presentation = biasUp * spikes(members) = %.0f\n", nm ? biasUp * sm / nm : 0.0);
        o.f("======================================================================\n");
    }

    // references: clamped throughout, and cue-then-released, mm:1080-1097
    std::vector<Vec> refClampCols(NSTIM), refFreeCols(NSTIM), refClampBal(NSTIM), refFreeBal(NSTIM);
    std::vector<Vec> asmSpkHost(NSTIM, Vec(NCOL, 0));
    for (uint32_t p = 0; p < NSTIM; ++p) {
        ora_present(n, p, 1.0, T, T, 0, 2000 + p, 2, nullptr);
        const uint32_t* s = final_spikes(T);
        refClampCols[p] = cols_of(s, IN, NCOL);
        refClampBal[p].assign(s, s + NCOL);
        asmSpkHost[p] = refClampBal[p];
        for (uint32_t col = 0; col < IN; ++col) asmSpkHost[p][col] = 0u;
        ora_present(n, p, 1.0, T, CLAMP, 0, 2100 + p, 2, nullptr);
        s = final_spikes(T);
        refFreeCols[p] = cols_of(s, IN, NCOL);
        refFreeBal[p].assign(s, s + NCOL);
    }
    double sz = 0;
    for (auto& a : refClampCols) sz += a.size();
    sz /= NSTIM;
    double xtalk = 0, xtalkF = 0;
    int xn = 0;
    for (uint32_t i = 0; i < NSTIM; ++i)
        for (uint32_t j = i + 1; j < NSTIM; ++j) {
            xtalk += overlap(refClampCols[i], refClampCols[j]);
            xtalkF += overlap(refFreeCols[i], refFreeCols[j]);
            xn++;
        }
    xtalk = xn ? xtalk / xn : 0;
    xtalkF = xn ? xtalkF / xn : 0;

    // STEP 0 diagnostic, mm:1107-1151
    if (c.diag) {
        Vec mp(N, 0), ap((size_t)NSTIM * NCOL, 0);
        for (uint32_t q = 0; q < NSTIM; ++q) std::copy(asmSpkHost[q].begin(), asmSpkHost[q].end(), ap.begin() + (size_t)q * NCOL);
        for (uint32_t col = IN; col < NCOL; ++col)
            for (uint32_t l = 0; l < 32; ++l) {
                uint32_t m = 0;
                for (uint32_t q = 0; q < std::min(NSTIM, 32u); ++q)
                    if ((asmSpkHost[q][col] >> l) & 1u) m |= 1u << q;
                mp[col * 32 + l] = m;
            }
        uint32_t nMember = 0;
        for (uint32_t i = 0; i < N; ++i)
            if (mp[i]) nMember++;
        const char* selName[2] = {"RECURRENT branches", "AFFERENT branches"};
        o.f("\n======================= STEP 0 DIAGNOSTIC =======================\n");
        o.f("4-bit synaptic counter distribution, %u stimuli, %u assembly-member neurons\n", NSTIM, nMember);
        for (uint32_t sel = 0; sel < 2; ++sel) {
            uint32_t h[48];
            // membership words hold 32 assemblies; past 32 the reference's shift (mm:619) is undefined
            ora_wdiag(n, mp.data(), ap.data(), std::min(NSTIM, 32u), sel, h);
            const char* cat[3] = {"within-assembly", "between-assembly", "all"};
            o.f("\n--- %s ---\n", selName[sel]);
            o.f("  %-17s %10s  %6s %6s %6s | %s\n", "pair type", "count", "mean", "%at 0", "%at 15", "histogram v=0..15");
            for (uint32_t cc = 0; cc < 3; ++cc) {
                double tot = 0, sum = 0;
                for (uint32_t v = 0; v < 16; ++v) { tot += h[cc * 16 + v]; sum += (double)v * h[cc * 16 + v]; }
                o.f("  %-17s %10.0f  %6.2f %6.1f %6.1f | ", cat[cc], tot, tot ? sum / tot : 0.0,
                    tot ? 100.0 * h[cc * 16 + 0] / tot : 0.0, tot ? 100.0 * h[cc * 16 + 15] / tot : 0.0);
                for (uint32_t v = 0; v < 16; ++v) o.f("%s%.1f", v ? "," : "", tot ? 100.0 * h[cc * 16 + v] / tot : 0.0);
                o.f("\n");
            }
        }
        o.f("================================================================\n");
    }

    if (!quiet) {  // mm:1153-1159 without the timing line
        o.f("  assembly size %.1f cols (%.2f%% of area) | cross-talk CLAMPED %.3f  FREE %.3f  (chance %.3f)\n\n", sz,
            100.0 * sz / (NCOL - IN), xtalk, xtalkF, sz / (NCOL - IN));
        o.f("--- pattern completion: cue %u steps, RELEASED %u.  BOTH reference definitions ---\n", CLAMP, T - CLAMP);
        o.f("  cue  |  CLAMPED ref: self  other  margin  |  FREE ref: self  other  margin\n");
    }

    // STEP A1: functional in-degree and winner identity, mm:1342-1409
    Vec ind(N, 0);
    ora_indeg(n, ind.data());
    {
        Vec v;
        for (uint32_t col = IN; col < NCOL; ++col)
            for (uint32_t l = 0; l < 32; ++l) v.push_back(ind[col * 32 + l]);
        Vec sv = v;
        std::sort(sv.begin(), sv.end());
        auto pct = [&](double q) { return sv[(size_t)(q * (sv.size() - 1))]; };
        double mean = 0;
        for (uint32_t x : v) mean += x;
        mean /= v.size();
        double sd = 0;
        for (uint32_t x : v) sd += (x - mean) * (x - mean);
        sd = std::sqrt(sd / v.size());
        o.f("\n=============================== STEP A1 ===============================\n");
        o.f("Functional in-degree over the %u RECURRENT branches (popcount of MSB plane)\n", 32 - inBranches);
        o.f("  n=%zu neurons   mean %.1f   sd %.1f   cv %.3f\n", v.size(), mean, sd, sd / mean);
        o.f("  min %u | p10 %u | p25 %u | p50 %u | p75 %u | p90 %u | p99 %u | p99.9 %u | max %u\n", sv.front(),
            pct(0.10), pct(0.25), pct(0.50), pct(0.75), pct(0.90), pct(0.99), pct(0.999), sv.back());
        o.f("  tail ratio p99/p50 = %.2f   max/p50 = %.2f\n", (double)pct(0.99) / pct(0.50),
            (double)sv.back() / pct(0.50));
        for (int which = 0; which < 2; ++which) {
            const auto& src = which ? refClampBal : refFreeBal;
            std::vector<Vec> ns(NSTIM);
            for (uint32_t q = 0; q < NSTIM; ++q) ns[q] = neuron_set(src[q].data(), IN, NCOL);
            double jm = 0, jmax = 0;
            int jn = 0;
            for (uint32_t i = 0; i < NSTIM; ++i)
                for (uint32_t j = i + 1; j < NSTIM; ++j) {
                    double x = jacc(ns[i], ns[j]);
                    jm += x; jn++; jmax = std::max(jmax, x);
                }
            Vec cnt(N, 0);
            for (auto& a : ns)
                for (uint32_t x : a) cnt[x]++;
            uint32_t nAll = 0, nNone = 0, nAny = 0;
            double dAll = 0, dNone = 0, dAny = 0, bAll = 0, bNone = 0, bAny = 0;
            for (uint32_t col = IN; col < NCOL; ++col)
                for (uint32_t l = 0; l < 32; ++l) {
                    uint32_t id = col * 32 + l;
                    if (cnt[id] >= NSTIM - 1) { nAll++; dAll += ind[id]; bAll += bias[id]; }
                    if (cnt[id] == 0) { nNone++; dNone += ind[id]; bNone += bias[id]; }
                    if (cnt[id] > 0) { nAny++; dAny += ind[id]; bAny += bias[id]; }
                }
            o.f("\n  --- winner identity, %s endpoint ---\n", which ? "CLAMPED" : "FREE-RUN");
            o.f("    mean pairwise Jaccard %.3f (max %.3f) over %d pairs, mean set size %.0f\n", jm / jn, jmax, jn,
                (double)ns[0].size());
            o.f("    neurons winning in >=%u/%u stimuli: %6u  in-degree %6.1f  intrinsic-bias %8.1f\n", NSTIM - 1,
                NSTIM, nAll, nAll ? dAll / nAll : 0.0, nAll ? bAll / nAll : 0.0);
            o.f("    neurons winning in any stimulus  : %6u  in-degree %6.1f  intrinsic-bias %8.1f\n", nAny,
                nAny ? dAny / nAny : 0.0, nAny ? bAny / nAny : 0.0);
            o.f("    neurons never winning            : %6u  in-degree %6.1f  intrinsic-bias %8.1f\n", nNone,
                nNone ? dNone / nNone : 0.0, nNone ? bNone / nNone : 0.0);
            o.f("    HUB RATIO (winners-in-all / never) = %.3f\n",
                (nNone && dNone) ? (dAll / std::max(1u, nAll)) / (dNone / nNone) : 0.0);
        }
        o.f("======================================================================\n");
    }

    // STEP A3 (chance baseline) + A4 (bias zeroed at readout), mm:1411-1511
    std::vector<Vec> refNBCols(NSTIM);
    {
        const double npool = (double)(NCOL - IN) * 32;
        std::vector<Vec> cns(NSTIM);
        for (uint32_t q = 0; q < NSTIM; ++q) cns[q] = neuron_set(refClampBal[q].data(), IN, NCOL);
        double msz = 0;
        for (auto& a : cns) msz += a.size();
        msz /= NSTIM;
        double eInt = msz * msz / npool;
        double eJac = eInt / (2.0 * msz - eInt);
        double jm = 0, jmax = 0, totInt = 0;
        int jn = 0;
        for (uint32_t i = 0; i < NSTIM; ++i)
            for (uint32_t j = i + 1; j < NSTIM; ++j) {
                double x = jacc(cns[i], cns[j]);
                jm += x; jn++; jmax = std::max(jmax, x);
                Vec ii;
                std::set_intersection(cns[i].begin(), cns[i].end(), cns[j].begin(), cns[j].end(),
                                      std::back_inserter(ii));
                totInt += ii.size();
            }
        Vec uni;
        for (auto& a : cns) uni.insert(uni.end(), a.begin(), a.end());
        std::sort(uni.begin(), uni.end());
        size_t uniqN = std::unique(uni.begin(), uni.end()) - uni.begin();
        double sumSz = msz * NSTIM;
        o.f("\n=============================== STEP A3 ===============================\n");
        o.f("Chance baseline for %u assemblies of %.0f neurons from a pool of %.0f\n", NSTIM, msz, npool);
        o.f("  E[|A n B|] = m^2/n            = %.2f neurons\n", eInt);
        o.f("  E[Jaccard] = E/(2m-E)         = %.5f      <-- CHANCE\n", eJac);
        o.f("  measured mean pairwise Jaccard = %.5f   (max %.5f)\n", jm / jn, jmax);
        o.f("  ratio measured/chance          = %.3f\n", eJac > 0 ? (jm / jn) / eJac : 0.0);
        o.f("  total pairwise overlap: measured %.0f neurons vs %.0f expected at chance (%.1fx less)\n", totInt,
            eInt * jn, totInt > 0 ? eInt * jn / totInt : 1e9);
        o.f("  sum of assembly sizes %.0f | union %zu | excess %.0f  -> %s\n", sumSz, uniqN, sumSz - (double)uniqN,
            (sumSz - (double)uniqN) < 0.02 * sumSz ? "NEAR-EXACT PARTITION" : "overlapping");
        pin->asm_jacc = jn ? jm / jn : 0.0;

        for (uint32_t q = 0; q < NSTIM; ++q) {
            ora_present(n, q, 1.0, T, T, 0, 2200 + q, 0, nullptr);
            refNBCols[q] = cols_of(final_spikes(T), IN, NCOL);
        }
        double preserve = 0;
        for (uint32_t q = 0; q < NSTIM; ++q) preserve += overlap(refNBCols[q], refClampCols[q]);
        preserve /= NSTIM;
        double xNB = 0;
        int xn2 = 0;
        for (uint32_t i = 0; i < NSTIM; ++i)
            for (uint32_t j = i + 1; j < NSTIM; ++j) { xNB += overlap(refNBCols[i], refNBCols[j]); xn2++; }
        xNB /= xn2;
        o.f("\n=============================== STEP A4 ===============================\n");
        o.f("Zero the intrinsic bias at TEST time only. Trained weights unchanged.\n");
        o.f("  clamped assembly preserved when bias is zeroed: overlap %.3f\n", preserve);
        o.f("  cross-talk of clamped assemblies, bias OFF     : %.3f  (bias ON: %.3f)\n", xNB, xtalk);
        o.f("\n  release trajectory, cue 75%%, bias OFF at test (vs bias ON baseline)\n");
        o.f("  step  phase   ov(clampRef,biasON)  ov(clampRef,biasOFF)  ov(noBiasRef,biasOFF)\n");
        std::vector<double> a1(T, 0), a2(T, 0), a3(T, 0);
        for (uint32_t q = 0; q < NSTIM; ++q) {
            ora_present(n, q, 0.75, T, CLAMP, 0, 4300 + q, 1, nullptr);
            for (uint32_t t = 0; t < T; ++t) a1[t] += overlap(cols_of(hist + (size_t)t * NCOL, IN, NCOL), refClampCols[q]);
            ora_present(n, q, 0.75, T, CLAMP, 0, 4300 + q, 0, nullptr);
            for (uint32_t t = 0; t < T; ++t) {
                Vec a = cols_of(hist + (size_t)t * NCOL, IN, NCOL);
                a2[t] += overlap(a, refClampCols[q]);
                a3[t] += overlap(a, refNBCols[q]);
            }
        }
        for (uint32_t t = 0; t < T; ++t)
            o.f("  %4u  %-5s %19.3f %21.3f %22.3f\n", t, t < CLAMP ? "cue" : "FREE", a1[t] / NSTIM, a2[t] / NSTIM,
                a3[t] / NSTIM);
        o.f("\n  pattern completion with bias OFF at test  (keep=0 is the no-cue control)\n");
        o.f("  cue  |  vs CLAMPED ref(biasON): self  other  margin | vs noBias ref: self  other  margin\n");
        for (double keep : {1.0, 0.75, 0.5, 0.35, 0.25, 0.15, 0.10, 0.05, 0.02, 0.0}) {
            double s1 = 0, o1 = 0, s2 = 0, o2 = 0;
            for (uint32_t q = 0; q < NSTIM; ++q) {
                ora_present(n, q, keep, T, CLAMP, 0, 4400 + q, 0, nullptr);
                Vec a = cols_of(final_spikes(T), IN, NCOL);
                s1 += overlap(a, refClampCols[q]);
                s2 += overlap(a, refNBCols[q]);
                double b1 = 0, b2 = 0;
                for (uint32_t r = 0; r < NSTIM; ++r)
                    if (r != q) {
                        b1 = std::max(b1, overlap(a, refClampCols[r]));
                        b2 = std::max(b2, overlap(a, refNBCols[r]));
                    }
                o1 += b1; o2 += b2;
            }
            o.f("  %3.0f%% |%25.3f %6.3f %+7.3f |%18.3f %6.3f %+7.3f\n", keep * 100, s1 / NSTIM, o1 / NSTIM,
                s1 / NSTIM - o1 / NSTIM, s2 / NSTIM, o2 / NSTIM, s2 / NSTIM - o2 / NSTIM);
            if (keep == 0.75) {
                pin->recall75 = s1 / NSTIM;
                pin->other75 = o1 / NSTIM;
                pin->margin75 = s1 / NSTIM - o1 / NSTIM;
            }
        }
        o.f("======================================================================\n");
    }

    // A4b (recalled subsets across cue levels) + A4c (bias growth), mm:1513-1599
    {
        Vec unt(N, 0);
        ora_indeg(n, ind.data());
        ora_untouched(n, unt.data());
        const double RECSYN = (double)(32 - inBranches) * 256;
        const double cues[4] = {1.0, 0.75, 0.35, 0.15};
        o.f("\n=============================== A4b ===============================\n");
        o.f("Is the recalled ~25%% a FIXED subset? (readout gate on: bias off at test)\n");
        std::vector<std::vector<Vec>> rec(NSTIM, std::vector<Vec>(4));
        for (uint32_t q = 0; q < NSTIM; ++q)
            for (int ci = 0; ci < 4; ++ci) {
                ora_present(n, q, cues[ci], T, CLAMP, 0, 5000 + q * 16 + ci, 2, nullptr);
                rec[q][ci] = neuron_set(final_spikes(T), IN, NCOL);
            }
        o.f("  pairwise Jaccard of recalled sets ACROSS cue levels, same stimulus:\n    ");
        double jall = 0;
        int jn = 0;
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                double m = 0;
                for (uint32_t q = 0; q < NSTIM; ++q) m += jacc(rec[q][a], rec[q][b]);
                m /= NSTIM;
                o.f("%.0f%%/%.0f%%=%.3f  ", cues[a] * 100, cues[b] * 100, m);
                jall += m; jn++;
            }
        o.f("\n    mean across cue pairs = %.3f\n", jall / jn);
        double coreSz = 0, periSz = 0, asmSz = 0, dCore = 0, dPeri = 0, uCore = 0, uPeri = 0;
        size_t nCore = 0, nPeri = 0;
        for (uint32_t q = 0; q < NSTIM; ++q) {
            Vec core = rec[q][0];
            for (int ci = 1; ci < 4; ++ci) {
                Vec t;
                std::set_intersection(core.begin(), core.end(), rec[q][ci].begin(), rec[q][ci].end(), std::back_inserter(t));
                core = t;
            }
            Vec A = neuron_set(refClampBal[q].data(), IN, NCOL), coreInA, peri;
            std::set_intersection(core.begin(), core.end(), A.begin(), A.end(), std::back_inserter(coreInA));
            std::set_difference(A.begin(), A.end(), core.begin(), core.end(), std::back_inserter(peri));
            coreSz += coreInA.size(); periSz += peri.size(); asmSz += A.size();
            for (uint32_t x : coreInA) { dCore += ind[x]; uCore += 100.0 * unt[x] / RECSYN; nCore++; }
            for (uint32_t x : peri) { dPeri += ind[x]; uPeri += 100.0 * unt[x] / RECSYN; nPeri++; }
        }
        o.f("\n  assembly %.0f neurons | CORE (recalled at all 4 cues) %.0f (%.1f%%) | periphery %.0f\n",
            asmSz / NSTIM, coreSz / NSTIM, 100.0 * coreSz / asmSz, periSz / NSTIM);
        o.f("  %-12s %10s %14s %22s\n", "group", "n", "in-degree", "%% synapses still at 5/6");
        o.f("  %-12s %10zu %14.1f %22.1f\n", "CORE", nCore, nCore ? dCore / nCore : 0.0, nCore ? uCore / nCore : 0.0);
        o.f("  %-12s %10zu %14.1f %22.1f\n", "periphery", nPeri, nPeri ? dPeri / nPeri : 0.0, nPeri ? uPeri / nPeri : 0.0);
        double allD = 0, allU = 0;
        size_t nAll = 0;
        for (uint32_t col = IN; col < NCOL; ++col)
            for (uint32_t l = 0; l < 32; ++l) {
                allD += ind[col * 32 + l]; allU += 100.0 * unt[col * 32 + l] / RECSYN; nAll++;
            }
        o.f("  %-12s %10zu %14.1f %22.1f\n", "all neurons", nAll, allD / nAll, allU / nAll);
        Vec bv;
        for (uint32_t col = IN; col < NCOL; ++col)
            for (uint32_t l = 0; l < 32; ++l) bv.push_back(bias[col * 32 + l]);
        std::sort(bv.begin(), bv.end());
        double bm = 0;
        for (uint32_t x : bv) bm += x;
        bm /= bv.size();
        o.f("\n=============================== A4c ===============================\n");
        o.f("Intrinsic bias growth after %u epochs (unbounded, non-leaky, learning-time only)\n", EPOCHS);
        o.f("  mean %.0f | p50 %u | p99 %u | max %u | UINT32_MAX %u | headroom %.3g x\n", bm, bv[bv.size() / 2],
            bv[(size_t)(0.99 * (bv.size() - 1))], bv.back(), 4294967295u, bv.back() ? 4294967295.0 / bv.back() : 0.0);
        o.f("  growth rate ~= %.1f per epoch (max/epochs) -> epochs to saturate ~ %.3g\n",
            EPOCHS ? (double)bv.back() / EPOCHS : 0.0,
            (bv.back() && EPOCHS) ? 4294967295.0 / ((double)bv.back() / EPOCHS) : 0.0);
        o.f("==================================================================\n");
    }

    const uint32_t T35 = std::min(35u, MAXT);
    // multiplicity of each neuron over the clamped references (bMult, mm:1604-1607 / 1656-1659)
    auto fill_mult = [&]() {
        std::fill(multBuf, multBuf + N, 0u);
        for (uint32_t q = 0; q < NSTIM; ++q)
            for (uint32_t x : neuron_set(refClampBal[q].data(), IN, NCOL)) multBuf[x]++;
    };

    // UNIONCHAR: the population visited during free running, mm:1601-1652
    if (c.unionchar) {
        fill_mult();
        const uint32_t NQ = std::min(NSTIM, 12u);
        double nIn = 0, nOut = 0, fI = 0, pI = 0, fO = 0, pO = 0, fA = 0, pA = 0, mOut = 0, mAll = 0, nAll = 0;
        Vec post(N, 0);
        Vec allRec;
        for (uint32_t col = IN; col < NCOL; ++col)
            for (uint32_t l = 0; l < 32; ++l) allRec.push_back(col * 32 + l);
        auto runW = [&](const Vec& sel, double& f, double& pp) {
            std::fill(post.begin(), post.end(), 0u);
            for (uint32_t x : sel) post[x] = 1u;
            uint32_t w[16];
            ora_within(n, refBall, post.data(), w);
            f += w[0]; pp += w[1];
        };
        for (uint32_t q = 0; q < NQ; ++q) {
            memcpy(refBall, refClampBal[q].data(), NCOL * 4);
            ora_present(n, q, 1.0, T35, CLAMP, 0, 9500 + q, 2, nullptr);
            Vec uni;
            for (uint32_t t = CLAMP; t < T35; ++t)
                for (uint32_t x : neuron_set(hist + (size_t)t * NCOL, IN, NCOL)) uni.push_back(x);
            std::sort(uni.begin(), uni.end());
            uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
            Vec ref = neuron_set(refClampBal[q].data(), IN, NCOL), in_, out_;
            std::set_intersection(uni.begin(), uni.end(), ref.begin(), ref.end(), std::back_inserter(in_));
            std::set_difference(uni.begin(), uni.end(), ref.begin(), ref.end(), std::back_inserter(out_));
            nIn += in_.size(); nOut += out_.size();
            runW(in_, fI, pI); runW(out_, fO, pO); runW(allRec, fA, pA);
            for (uint32_t x : out_) mOut += multBuf[x];
            for (uint32_t x : allRec) mAll += multBuf[x];
            nAll += allRec.size();
        }
        o.f("\nUNIONCHAR %.0f %.0f %.4f %.4f %.4f %.4f %.4f\n", nIn / NQ, nOut / NQ, 100.0 * fI / pI, 100.0 * fO / pO,
            100.0 * fA / pA, nOut > 0 ? mOut / nOut : 0.0, nAll > 0 ? mAll / nAll : 0.0);
    }

    // SHARED vs EXCLUSIVE members, mm:1654-1692
    if (c.split) {
        fill_mult();
        const uint32_t NQ = std::min(NSTIM, 12u);
        double fS = 0, pS = 0, fE = 0, pE = 0, nS = 0, nE = 0;
        std::vector<double> dS(64, 0), cS(64, 0), dE(64, 0), cE(64, 0);
        for (uint32_t q = 0; q < NQ; ++q) {
            memcpy(refBall, refClampBal[q].data(), NCOL * 4);
            uint32_t w[16];
            ora_within(n, refBall, nullptr, w);  // useMask = 0 (mm:1667-1668)
            fS += w[4]; pS += w[5]; fE += w[6]; pE += w[7]; nS += w[8]; nE += w[9];
            memset(dl, 0, ndl * 4);
            ora_present(n, q, 1.0, T35, CLAMP, 0, 8800 + q, 2, nullptr);
            for (uint32_t t = 0; t < T35; ++t) {
                dS[t] += dl[t * 4 + 0]; cS[t] += dl[t * 4 + 1];
                dE[t] += dl[t * 4 + 2]; cE[t] += dl[t * 4 + 3];
            }
        }
        o.f("\n===== SHARED vs EXCLUSIVE assembly members (NSTIM=%u) =====\n", NSTIM);
        o.f("  shared fraction of assembly : %.3f  (%.0f shared / %.0f exclusive neurons)\n",
            (nS + nE) > 0 ? nS / (nS + nE) : 0.0, nS / NQ, nE / NQ);
        o.f("  within-assembly density  SHARED    %.4f%%  (%.0f/%.0f)\n", 100.0 * fS / pS, fS, pS);
        o.f("  within-assembly density  EXCLUSIVE %.4f%%  (%.0f/%.0f)\n", 100.0 * fE / pE, fE, pE);
        o.f("  free-phase drive  SHARED   :");
        for (uint32_t t = CLAMP; t < T35; ++t) o.f(" %.0f", cS[t] ? dS[t] / cS[t] : 0.0);
        o.f("\n  free-phase drive  EXCLUSIVE:");
        for (uint32_t t = CLAMP; t < T35; ++t) o.f(" %.0f", cE[t] ? dE[t] / cE[t] : 0.0);
        o.f("\n==========================================================\n");
    }

    // within-assembly recurrent scaffold and drive trace, mm:1694-1730
    if (c.deep) {
        const uint32_t NQ = std::min(NSTIM, 12u);
        const uint32_t TT = std::min(std::min(T + 20u, 40u), MAXT);
        double fIn = 0, pIn = 0, fOut = 0, pOut = 0;
        std::vector<double> drvSum(40, 0), drvCnt(40, 0);
        for (uint32_t q = 0; q < NQ; ++q) {
            memcpy(refBall, refClampBal[q].data(), NCOL * 4);
            uint32_t w[16] = {0};
            ora_within(n, refBall, nullptr, w);
            fIn += w[0]; pIn += w[1]; fOut += w[2]; pOut += w[3];
            memset(dl, 0, ndl * 4);
            ora_present(n, q, 1.0, TT, CLAMP, 0, 8600 + q, 2, nullptr);
            for (uint32_t t = 0; t < TT; ++t) {
                drvSum[t] += dl[t * 4 + 0] + dl[t * 4 + 2];
                drvCnt[t] += dl[t * 4 + 1] + dl[t * 4 + 3];
            }
        }
        o.f("\n========= WITHIN-ASSEMBLY RECURRENT SCAFFOLD (NSTIM=%u) =========\n", NSTIM);
        o.f("  within-assembly : %.0f functional of %.0f possible = %.4f%%\n", fIn, pIn, 100.0 * fIn / pIn);
        o.f("  between         : %.0f functional of %.0f possible = %.4f%%\n", fOut, pOut, 100.0 * fOut / pOut);
        o.f("  enrichment within/between = %.2fx\n", (fOut > 0 && pIn > 0) ? (fIn / pIn) / (fOut / pOut) : 0.0);
        o.f("  free-phase drive to assembly members, per step:\n   ");
        for (uint32_t t = 0; t < TT; ++t)
            if (t < CLAMP + 12) o.f(" t%u:%.0f", t, drvCnt[t] ? drvSum[t] / drvCnt[t] : 0.0);
        o.f("\n================================================================\n");
        pin->fin = fIn; pin->pin = pIn; pin->fout = fOut; pin->pout = pOut;
        pin->within_pct = 100.0 * fIn / pIn;
        pin->between_pct = 100.0 * fOut / pOut;
        pin->drive_tt = TT;
        for (uint32_t t = 0; t < TT; ++t) pin->drive_t[t] = drvCnt[t] ? drvSum[t] / drvCnt[t] : 0.0;
    }

    // multiplicity of the retrieved population and the heavy tail, mm:1732-1801
    if (c.deep) {
        std::vector<uint16_t> mlt(N, 0);
        for (uint32_t q = 0; q < NSTIM; ++q)
            for (uint32_t x : neuron_set(refClampBal[q].data(), IN, NCOL)) mlt[x]++;
        double popSum = 0;
        uint32_t popN = 0, maxMult = 0;
        for (uint32_t i = IN * 32; i < N; ++i)
            if (mlt[i]) { popSum += mlt[i]; popN++; maxMult = std::max<uint32_t>(maxMult, mlt[i]); }
        const double popMean = popN ? popSum / popN : 0.0;
        const uint32_t NQ = std::min(NSTIM, 30u);
        double mIn = 0, mOut = 0, nIn = 0, nOut = 0, sumBest = 0, sumMeanOther = 0, sumSelf = 0;
        for (uint32_t q = 0; q < NQ; ++q) {
            ora_present(n, q, 1.0, T, CLAMP, 0, 8100 + q, 2, nullptr);
            const uint32_t* sp = final_spikes(T);
            Vec recN = neuron_set(sp, IN, NCOL), refN = neuron_set(refClampBal[q].data(), IN, NCOL), in_, out_;
            std::set_intersection(recN.begin(), recN.end(), refN.begin(), refN.end(), std::back_inserter(in_));
            std::set_difference(recN.begin(), recN.end(), refN.begin(), refN.end(), std::back_inserter(out_));
            for (uint32_t x : in_) { mIn += mlt[x]; nIn++; }
            for (uint32_t x : out_) { mOut += mlt[x]; nOut++; }
            Vec recCols = cols_of(sp, IN, NCOL);
            sumSelf += overlap(recCols, refClampCols[q]);
            double best = 0, tot = 0;
            uint32_t cnt = 0;
            for (uint32_t r = 0; r < NSTIM; ++r)
                if (r != q) {
                    double ov = overlap(recCols, refClampCols[r]);
                    best = std::max(best, ov); tot += ov; cnt++;
                }
            sumBest += best;
            sumMeanOther += cnt ? tot / cnt : 0.0;
        }
        o.f("\n============ RETRIEVED-POPULATION MULTIPLICITY (NSTIM=%u) ============\n", NSTIM);
        o.f("  population: %u neurons used, mean multiplicity %.2f, max %u\n", popN, popMean, maxMult);
        o.f("  recalled AND in reference    : n=%.0f  mean multiplicity %.2f\n", nIn, nIn ? mIn / nIn : 0.0);
        o.f("  recalled but NOT in reference: n=%.0f  mean multiplicity %.2f  (%.2fx population)\n", nOut,
            nOut ? mOut / nOut : 0.0, (nOut && popMean) ? (mOut / nOut) / popMean : 0.0);
        o.f("  self %.3f | best-other %.3f | MEAN-over-others %.3f  -> %s\n", sumSelf / NQ, sumBest / NQ,
            sumMeanOther / NQ, (sumMeanOther / NQ > 0.5 * sumBest / NQ) ? "UNIFORM contamination" : "TAIL-dominated");
        std::vector<Vec> ns(NSTIM);
        for (uint32_t q = 0; q < NSTIM; ++q) ns[q] = neuron_set(refClampBal[q].data(), IN, NCOL);
        std::vector<double> J, IO, IX;
        for (uint32_t i = 0; i < NSTIM; ++i)
            for (uint32_t j = i + 1; j < NSTIM; ++j) {
                J.push_back(jacc_union(ns[i], ns[j]));
                Vec a = stimCols[i], b = stimCols[j];
                std::sort(a.begin(), a.end());
                std::sort(b.begin(), b.end());
                Vec ii;
                std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(ii));
                IO.push_back((double)ii.size());
                IX.push_back((double)std::max(i, j));
            }
        if (!J.empty()) {  // the reference indexes J unguarded (mm:1799); NSTIM >= 2 there
            std::vector<double> Js = J;
            std::sort(Js.begin(), Js.end());
            o.f("  heavy tail: %zu pairs | mean J %.5f | p99 %.5f | max %.5f\n", J.size(),
                std::accumulate(J.begin(), J.end(), 0.0) / J.size(), Js[(size_t)(0.99 * (Js.size() - 1))], Js.back());
            o.f("  corr(assemblyJaccard, INPUT overlap)      r = %+.3f\n", pearson(J, IO));
            o.f("  corr(assemblyJaccard, allocation index)   r = %+.3f\n", pearson(J, IX));
        }
        o.f("======================================================================\n");
    }

    // long free run: PARITY line and WAVE table, mm:1803-1858
    if (c.wave) {
        const uint32_t TW = std::min(std::min(c.wavet, 60u), MAXT);
        const uint32_t NQ = std::min(NSTIM, 12u);
        std::vector<double> ovR(TW, 0), nCol(TW, 0), nNeu(TW, 0), fracA(TW, 0), jPrev(TW, 0), jPrev2(TW, 0);
        double unionSz = 0, unionE = 0, unionO = 0, unionX = 0;
        for (uint32_t q = 0; q < NQ; ++q) {
            ora_present(n, q, 1.0, TW, CLAMP, 0, 7000 + q, 2, nullptr);
            Vec refN = neuron_set(refClampBal[q].data(), IN, NCOL);
            Vec prev, prev2, uni, uniE, uniO;
            for (uint32_t t = 0; t < TW; ++t) {
                const uint32_t* row = hist + (size_t)t * NCOL;
                Vec cur = neuron_set(row, IN, NCOL);
                uint32_t nc = 0;
                for (uint32_t col = IN; col < NCOL; ++col)
                    if (row[col]) nc++;
                nCol[t] += nc; nNeu[t] += cur.size();
                Vec ii;
                std::set_intersection(cur.begin(), cur.end(), refN.begin(), refN.end(), std::back_inserter(ii));
                ovR[t] += (cur.empty() || refN.empty()) ? 0.0
                                                        : (double)ii.size() / std::sqrt((double)cur.size() * refN.size());
                fracA[t] += refN.empty() ? 0.0 : (double)ii.size() / refN.size();
                if (t >= CLAMP) {
                    uni.insert(uni.end(), cur.begin(), cur.end());
                    if (t % 2u == 0u) uniE.insert(uniE.end(), cur.begin(), cur.end());
                    else uniO.insert(uniO.end(), cur.begin(), cur.end());
                }
                jPrev[t] += t ? jacc(cur, prev) : 0.0;
                jPrev2[t] += t > 1 ? jacc(cur, prev2) : 0.0;
                prev2 = prev; prev = cur;
            }
            std::sort(uni.begin(), uni.end());
            uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
            unionSz += (double)uni.size();
            std::sort(uniE.begin(), uniE.end());
            uniE.erase(std::unique(uniE.begin(), uniE.end()), uniE.end());
            std::sort(uniO.begin(), uniO.end());
            uniO.erase(std::unique(uniO.begin(), uniO.end()), uniO.end());
            Vec ii;
            std::set_intersection(uniE.begin(), uniE.end(), uniO.begin(), uniO.end(), std::back_inserter(ii));
            unionE += (double)uniE.size();
            unionO += (double)uniO.size();
            unionX += (uniE.empty() || uniO.empty()) ? 0.0 : (double)ii.size() / std::sqrt((double)uniE.size() * uniO.size());
        }
        o.f("\nPARITY %.0f %.4f %.0f %.4f %.0f %.4f %.4f\n", unionSz / NQ, unionSz / NQ / 1536.0, unionE / NQ,
            unionE / NQ / 1536.0, unionO / NQ, unionO / NQ / 1536.0, unionX / NQ);
        o.f("\n================= WAVE: %u free steps after a %u-step cue =================\n", TW - CLAMP, CLAMP);
        o.f("  step  phase   cols  neurons   ov(clampRef)  fracOfAssembly  J(t,t-1)   J(t,t-2)\n");
        for (uint32_t t = 0; t < TW; ++t)
            o.f("  %4u  %-5s %6.0f %8.0f %13.3f %15.3f %10.3f %10.3f\n", t, t < CLAMP ? "cue" : "FREE", nCol[t] / NQ,
                nNeu[t] / NQ, ovR[t] / NQ, fracA[t] / NQ, jPrev[t] / NQ, jPrev2[t] / NQ);
        o.f("==========================================================================\n");
        pin->wave_tw = TW;
        for (uint32_t t = 0; t < TW; ++t) pin->wave_ov[t] = ovR[t] / NQ;
        if (TW > CLAMP) {
            // bench/cap10.sh:19-24: ratio of the printed 3-decimal overlaps
            char a[64], l[64];
            snprintf(a, sizeof a, "%.3f", ovR[CLAMP] / NQ);
            snprintf(l, sizeof l, "%.3f", ovR[TW - 1] / NQ);
            double av = atof(a), lv = atof(l);
            pin->sustained = av != 0.0 ? lv / av : 0.0;
        }
    }

    // free-run diagnostic at a 75% cue, mm:1860-1887
    {
        o.f("\n============ FREE-RUN DIAGNOSTIC (cue %u steps, then released) ============\n", CLAMP);
        o.f("  cue=75%%, averaged over %u stimuli.  'margin' = k-th minus (k+1)-th drive\n", NSTIM);
        o.f("  step  phase  cols  drive_k  margin  %%marg=0   ov(CLAMPED ref)  ov(FREE ref)\n");
        std::vector<double> ovt(T, 0), ovf(T, 0), mg(T, 0), dk(T, 0), tie(T, 0), zer(T, 0), nc(T, 0);
        for (uint32_t q = 0; q < NSTIM; ++q) {
            ora_present(n, q, 0.75, T, CLAMP, 0, 4000 + q, 2, nullptr);
            for (uint32_t t = 0; t < T; ++t) {
                Vec a = cols_of(hist + (size_t)t * NCOL, IN, NCOL);
                ovt[t] += overlap(a, refClampCols[q]);
                ovf[t] += overlap(a, refFreeCols[q]);
                const uint32_t* sb = stats + 64 + t * 8;
                double cnt = sb[2];
                nc[t] += cnt;
                if (cnt > 0) {
                    mg[t] += sb[0] / cnt; dk[t] += sb[1] / cnt;
                    tie[t] += 100.0 * sb[3] / cnt; zer[t] += 100.0 * sb[4] / cnt;
                }
            }
        }
        for (uint32_t t = 0; t < T; ++t)
            o.f("  %4u  %-5s %5.0f %8.1f %7.1f %8.1f   %14.3f  %12.3f\n", t, t < CLAMP ? "cue" : "FREE", nc[t] / NSTIM,
                dk[t] / NSTIM, mg[t] / NSTIM, tie[t] / NSTIM, ovt[t] / NSTIM, ovf[t] / NSTIM);
        o.f("==========================================================================\n");
    }

    // completion against both references, mm:1889-1906
    for (double keep : {1.0, 0.75, 0.5, 0.35, 0.25}) {
        double self = 0, other = 0, selfF = 0, otherF = 0;
        for (uint32_t p = 0; p < NSTIM; ++p) {
            ora_present(n, p, keep, T, CLAMP, 0, 3000 + p, 2, nullptr);
            Vec a = cols_of(final_spikes(T), IN, NCOL);
            self += overlap(a, refClampCols[p]);
            selfF += overlap(a, refFreeCols[p]);
            double best = 0, bestF = 0;
            for (uint32_t q = 0; q < NSTIM; ++q)
                if (q != p) {
                    best = std::max(best, overlap(a, refClampCols[q]));
                    bestF = std::max(bestF, overlap(a, refFreeCols[q]));
                }
            other += best; otherF += bestF;
        }
        self /= NSTIM; other /= NSTIM; selfF /= NSTIM; otherF /= NSTIM;
        if (!quiet)
            o.f("  %3.0f%% |  %16.3f %6.3f %+7.3f  |  %11.3f %6.3f %+7.3f\n", keep * 100, self, other, self - other,
                selfF, otherF, selfF - otherF);
    }
    if (!quiet) o.f("\n=== DONE ===\n");
    pin->wstat_final = ora_wstat(n);
    ora_net_destroy(n);
    if (buf && buf_len) snprintf(buf, buf_len, "%s", o.s.c_str());
    return 0;
}
