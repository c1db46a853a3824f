// AXON CPU oracle protocol -- TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Restates the state-mutating part of the reference main() (mm:996-1105:
// init, epochs x patterns learning presentations, the extra learning
// presentation of mm:1033) and the readouts that the paper pins:
//   references          mm:1084-1097
//   A3 Jaccard          mm:1425-1455
//   A4 completion       mm:1458-1509
//   deep density/drive  mm:1695-1730
//   wave                mm:1804-1858
// Non-learning presentations mutate no persistent state, so the settle loop
// (mm:1001) and the read-only diagnostic blocks are not needed to reach the
// pinned numbers and are skipped here.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"

namespace {

struct Out {
    std::string s;
    void f(const char* fmt, ...) __attribute__((format(printf, 2, 3))) {
        char b[4096];
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
double jacc(const Vec& a, const Vec& b) {
    if (a.empty() || b.empty()) return 0.0;
    Vec i, u;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(i));
    std::set_union(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(u));
    return (double)i.size() / u.size();
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

}  // namespace

extern "C" int ora_run_protocol(const ora_config* cin, int nthreads, ora_pinned* pin, char* buf,
                                size_t buf_len) {
    ora_config c = *cin;
    Out o;
    ora_net* n = ora_net_create(&c, nthreads);
    if (!n) return 1;
    const uint32_t NCOL = c.ncol, IN = c.incols, NSTIM = c.stim, T = c.T, CLAMP = c.clampsteps;
    const uint32_t trainClamp = c.trainclamp;
    memset(pin, 0, sizeof(*pin));
    size_t nw = 0;
    uint32_t* stats = ora_buffer(n, ORA_STATS, &nw);
    uint32_t* hist = ora_buffer(n, ORA_HIST, &nw);
    const double NSYN = (double)NCOL * 32.0 * 8192.0;

    for (uint32_t ep = 0; ep < c.epochs; ++ep) {
        for (uint32_t p = 0; p < NSTIM; ++p) ora_present(n, p, 1.0, T, trainClamp, 1, 1000 + p, 2, nullptr);
        if (!c.quiet && (ep % 5 == 4 || ep == 0)) {
            double ltp = stats[8], ltd = stats[9];
            uint32_t tp = stats[10], td = stats[11];
            double wd = 100.0 * ora_wstat(n) / NSYN;
            o.f("  ep %3u  wdens %7.3f%%  LTP %10.0f  LTD %10.0f  LTD/LTP %6.3f  learners %u/%u\n", ep + 1, wd,
                ltp, ltd, ltp > 0 ? ltd / ltp : 0.0, tp, td);
        }
    }
    ora_present(n, 0, 1.0, T, trainClamp, 1, 31337, 2, nullptr);  // mm:1033

    std::vector<Vec> refClampCols(NSTIM), refFreeCols(NSTIM), refClampBal(NSTIM), refNBCols(NSTIM);
    for (uint32_t p = 0; p < NSTIM; ++p) {
        ora_present(n, p, 1.0, T, T, 0, 2000 + p, 2, nullptr);
        const uint32_t* s = ora_final_spikes(n, T);
        refClampCols[p] = cols_of(s, IN, NCOL);
        refClampBal[p].assign(s, s + NCOL);
        ora_present(n, p, 1.0, T, CLAMP, 0, 2100 + p, 2, nullptr);
        refFreeCols[p] = cols_of(ora_final_spikes(n, T), IN, NCOL);
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
    o.f("  assembly size %.1f cols (%.2f%% of area) | cross-talk CLAMPED %.3f  FREE %.3f  (chance %.3f)\n\n", sz,
        100.0 * sz / (NCOL - IN), xtalk, xtalkF, sz / (NCOL - IN));

    // A3, mm:1425-1455
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
        o.f("  E[Jaccard] = E/(2m-E)         = %.5f      <-- CHANCE\n", eJac);
        o.f("  measured mean pairwise Jaccard = %.5f   (max %.5f)\n", jm / jn, jmax);
        o.f("  ratio measured/chance          = %.3f\n", eJac > 0 ? (jm / jn) / eJac : 0.0);
        o.f("  sum of assembly sizes %.0f | union %zu | excess %.0f  -> %s\n", sumSz, uniqN, sumSz - (double)uniqN,
            (sumSz - (double)uniqN) < 0.02 * sumSz ? "NEAR-EXACT PARTITION" : "overlapping");
        pin->asm_jacc = jn ? jm / jn : 0.0;
    }

    // A4 completion, mm:1458-1509
    for (uint32_t q = 0; q < NSTIM; ++q) {
        ora_present(n, q, 1.0, T, T, 0, 2200 + q, 0, nullptr);
        refNBCols[q] = cols_of(ora_final_spikes(n, T), IN, NCOL);
    }
    o.f("\n  pattern completion with bias OFF at test  (keep=0 is the no-cue control)\n");
    o.f("  cue  |  vs CLAMPED ref(biasON): self  other  margin | vs noBias ref: self  other  margin\n");
    for (double keep : {1.0, 0.75, 0.5, 0.35, 0.25, 0.15, 0.10, 0.05, 0.02, 0.0}) {
        double s1 = 0, o1 = 0, s2 = 0, o2 = 0;
        for (uint32_t q = 0; q < NSTIM; ++q) {
            ora_present(n, q, keep, T, CLAMP, 0, 4400 + q, 0, nullptr);
            Vec a = cols_of(ora_final_spikes(n, T), IN, NCOL);
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

    // deep: within-assembly density + drive trace, mm:1695-1730
    if (c.deep) {
        uint32_t NQ = std::min(NSTIM, 12u);
        double fIn = 0, pIn = 0, fOut = 0, pOut = 0;
        std::vector<double> drvSum(40, 0), drvCnt(40, 0);
        uint32_t TT = std::min(T + 20u, 40u);
        uint32_t* refBall = ora_buffer(n, ORA_REFBALL, &nw);
        uint32_t* dl = ora_buffer(n, ORA_DRIVELOG, &nw);
        size_t ndl = nw;
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

    // wave, mm:1804-1858 (ov(clampRef) column; the other columns are host-side)
    if (c.wave) {
        uint32_t TW = std::min(c.wavet, 60u);
        uint32_t NQ = std::min(NSTIM, 12u);
        std::vector<double> ovR(TW, 0), nCol(TW, 0), nNeu(TW, 0), fracA(TW, 0), jPrev(TW, 0), jPrev2(TW, 0);
        for (uint32_t q = 0; q < NQ; ++q) {
            ora_present(n, q, 1.0, TW, CLAMP, 0, 7000 + q, 2, nullptr);
            Vec refN = neuron_set(refClampBal[q].data(), IN, NCOL);
            Vec prev, prev2;
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
                jPrev[t] += t ? jacc(cur, prev) : 0.0;
                jPrev2[t] += t > 1 ? jacc(cur, prev2) : 0.0;
                prev2 = prev; prev = cur;
            }
        }
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
    pin->wstat_final = ora_wstat(n);
    ora_net_destroy(n);
    if (buf && buf_len) snprintf(buf, buf_len, "%s", o.s.c_str());
    return 0;
}
