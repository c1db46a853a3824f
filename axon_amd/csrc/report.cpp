// axon_run -- the reference program's whole run on the engine.
//
// A restatement of the reference's host main() (src/axon.mm, mm:669-1911)
// over libaxon's C ABI: the same key=value arguments (last occurrence wins),
// the same protocol order and the same report lines, so the reference's
// bench/*.sh scripts run unchanged against the CLI (axon_amd/csrc/cli.cpp,
// which calls this).  Kept in the library so that a caller holding networks
// trained elsewhere (axon_multi_train) can run the rest of the protocol and
// the report on them in-process.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/axon_api.h"

namespace axon {
// the calling thread's axon_last_error() message (kept by the C-ABI layer)
axon_status set_error(axon_status s, const std::string& msg);
}  // namespace axon

namespace {

using Vec = std::vector<uint32_t>;
constexpr uint32_t NPC = 32, NSLOT = 32, PRESYN = 256, FANIN = 8192;

// A failed engine call ends the run: axon_run returns its status.
struct RunError {
    axon_status s;
    std::string msg;
};
[[noreturn]] void die(const char* what, axon_status s) {
    throw RunError{s, std::string(what) + " failed: " + axon_status_str(s) + " (" + axon_last_error() + ")"};
}
#define CHECK(call)                                  \
    do {                                             \
        axon_status s_ = (call);                     \
        if (s_ != AXON_OK) die(#call, s_);           \
    } while (0)

// The report stream of the run on this thread.
thread_local FILE* g_out = nullptr;
void say(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void say(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vfprintf(g_out, fmt, ap);
    va_end(ap);
}

struct Engine {
    axon_net* n = nullptr;
    bool owned = true;  // false: the caller's trained network, left to the caller
    ~Engine() {
        if (n && owned) axon_net_destroy(n);
    }
    axon_config c{};
    uint32_t NCOL = 0, N = 0, IN = 0;
    Vec spk, hist, stats, dlog;
    uint32_t theta[2] = {0, 0};

    // present(), mm:815-888; outputs are copied to host like the reference's Shared buffers
    double present(uint32_t p, double keep, uint32_t T, uint32_t clampSteps, bool learn, bool freeze,
                   uint32_t seed, uint32_t useBias = 2u, bool outputs = true) {
        axon_present_args a{};
        a.pattern = p; a.keep = keep; a.T = T; a.clamp_steps = clampSteps; a.learn = learn; a.freeze = freeze;
        a.seed = seed; a.use_bias = useBias;
        axon_present_out o{};
        if (outputs) {
            o.spikes = spk.data(); o.hist = hist.data(); o.stats = stats.data(); o.drive_log = dlog.data();
            o.theta = theta;
        } else {
            a.flags = AXON_PRESENT_ASYNC;
        }
        CHECK(axon_present(n, &a, &o));
        return o.ms_per_step;
    }
    const uint32_t* row(uint32_t t) const { return hist.data() + (size_t)t * NCOL; }
    Vec cols_of(const uint32_t* s) const {
        Vec v;
        for (uint32_t c = IN; c < NCOL; ++c)
            if (s[c]) v.push_back(c);
        return v;
    }
    Vec assemblyOf() const { return cols_of(spk.data()); }  // mm:890-895
    Vec nset(const uint32_t* bal) const {
        Vec r;
        for (uint32_t c = IN; c < NCOL; ++c) {
            uint32_t b = bal[c];
            while (b) { uint32_t l = __builtin_ctz(b); b &= b - 1; r.push_back(c * NPC + l); }
        }
        return r;
    }
    double wstat() {  // mm:784-791
        uint32_t f = 0;
        CHECK(axon_wstat(n, &f));
        return 100.0 * f / ((double)N * FANIN);
    }
};

// String-valued keys of this CLI (no reference counterpart): prefix + '=',
// the last occurrence wins, as argu() does for the numeric keys (mm:660-667).
const char* arg_str(int argc, const char* const* argv, const char* key) {
    const char* v = nullptr;
    const size_t kl = strlen(key);
    for (int i = 1; i < argc; ++i)
        if (!strncmp(argv[i], key, kl) && argv[i][kl] == '=') v = argv[i] + kl + 1;
    return v;
}

double overlap(const Vec& a, const Vec& b) {  // mm:896-901
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

}  // namespace

// The keys that select readouts or printing only (the engine never reads
// them): a trained network may come from a run that set them otherwise.
static axon_config training_keys(axon_config x) {
    x.quiet = x.split = x.unionchar = x.diag = x.profile = x.vtopk = x.vbs = 0;
    x.deep = x.wave = x.wavet = 0;
    return x;
}

static void run(int argc, const char* const* argv, axon_net* trained) {
    Engine E;
    axon_config& c = E.c;
    CHECK(axon_config_parse(argc, argv, &c));
    CHECK(axon_config_validate(&c));
    // save=<path>: snapshot of the network after its training epochs (before
    // the extra presentation).  load=<path>: start from such a snapshot (made
    // with the same training keys) instead of running the epochs; everything
    // after them -- the extra presentation, references and readouts -- is
    // unchanged.  A `trained` network from the caller takes the place of a
    // loaded one.
    const char* savePath = arg_str(argc, argv, "save");
    const char* loadPath = arg_str(argc, argv, "load");
    const bool skipEpochs = loadPath || trained;
    const uint32_t NCOL = c.ncol, N = NCOL * NPC, IN = c.incols, NSTIM = c.stim, T = c.T, CLAMP = c.clampsteps;
    const uint32_t K = c.k, targetC = c.target, biasUp = c.biasup, inBranches = c.inbr, EPOCHS = c.epochs;
    const uint32_t trainClamp = c.trainclamp, quiet = c.quiet;
    const uint32_t NGRP = NCOL / 8;
    const double PLANE_B = (double)NCOL * NSLOT * 2 * 32 * 16;
    E.NCOL = NCOL; E.N = N; E.IN = IN;

    if (!quiet) {  // mm:737-747
        say("======================================================================\n");
        say(" AXON v0  -  columnar assembly engine, AMD Instinct MI355X (gfx950)\n");
        say("======================================================================\n");
        say("  neurons %u (%u cols x %u) | branches %u x %u presyn | fan-in %u | %.2f G syn\n", N, NCOL, NPC,
               NSLOT, PRESYN, FANIN, (double)N * FANIN / 1e9);
        say("  weights 4 planes x %.0f MB = %.0f MB | groups %u\n", PLANE_B / 1048576.0, 4 * PLANE_B / 1048576.0,
               NGRP);
        say("  nand=%u target=%u k=%u dthr=%u mode=%u epochs=%u T=%u clamp=%u stim=%u biasUp=%u inbr=%u "
               "stimcols=%u\n\n",
               c.nand, targetC, K, c.dthr, c.mode, EPOCHS, T, CLAMP, NSTIM, biasUp, inBranches, c.stimcols);
    }
    CHECK(axon_net_create(&c, &E.n));
    size_t w = 0;
    CHECK(axon_buffer_words(E.n, AXON_BUF_STATS, &w));
    E.stats.assign(w, 0);
    CHECK(axon_buffer_words(E.n, AXON_BUF_DRIVELOG, &w));
    E.dlog.assign(w, 0);
    CHECK(axon_buffer_words(E.n, AXON_BUF_HIST, &w));
    E.hist.assign(w, 0);
    E.spk.assign(NCOL, 0);
    const uint32_t MAXT = c.max_t;
    std::vector<Vec> stimCols(NSTIM, Vec(c.stimcols));
    for (uint32_t p = 0; p < NSTIM; ++p) CHECK(axon_stim_cols(E.n, p, stimCols[p].data()));

    if (c.vbs) {  // mm:903-956
        say("\n========== BIT-SERIAL VERIFICATION vs SCALAR REFERENCE ==========\n");
        const uint32_t NW = 1 << 14;
        std::mt19937 r(424242);
        struct Tst { const char* n; uint32_t op, lo, hi; };
        const Tst tests[] = {{"bs_inc (carry, saturate 15)", 0, 0, 0}, {"bs_dec (BORROW, saturate 0)", 1, 0, 0},
                             {"bs_clip [5,11]", 2, 5, 11}, {"bs_clip [6,10]", 2, 6, 10}, {"bs_clip [4,12]", 2, 4, 12}};
        for (const Tst& t : tests) {
            const size_t nbits = (size_t)NW * 4 * 32, nw = (size_t)NW * 4;
            std::vector<uint8_t> vin(nbits), mk(nbits);
            Vec p0(nw, 0), p1(nw, 0), p2(nw, 0), p3(nw, 0), pm(nw, 0);
            for (size_t i = 0; i < nbits; ++i) {
                uint8_t v = r() & 15u, m = r() & 1u;
                vin[i] = v; mk[i] = m;
                size_t wi = i >> 5, b = i & 31;
                if (v & 1) p0[wi] |= 1u << b;
                if (v & 2) p1[wi] |= 1u << b;
                if (v & 4) p2[wi] |= 1u << b;
                if (v & 8) p3[wi] |= 1u << b;
                if (m) pm[wi] |= 1u << b;
            }
            CHECK(axon_bs_test(t.op, t.lo, t.hi, p0.data(), p1.data(), p2.data(), p3.data(), pm.data(), nw));
            size_t bad = 0, firstBad = SIZE_MAX;
            std::vector<uint8_t> vout(nbits);
            for (size_t i = 0; i < nbits; ++i) {
                size_t wi = i >> 5, b = i & 31;
                vout[i] = ((p0[wi] >> b) & 1) | (((p1[wi] >> b) & 1) << 1) | (((p2[wi] >> b) & 1) << 2) |
                          (((p3[wi] >> b) & 1) << 3);
                int exp;
                if (t.op == 0) exp = mk[i] ? std::min(15, (int)vin[i] + 1) : vin[i];
                else if (t.op == 1) exp = mk[i] ? std::max(0, (int)vin[i] - 1) : vin[i];
                else exp = std::min((int)t.hi, std::max((int)t.lo, (int)vin[i]));
                if (exp != (int)vout[i]) { if (bad == 0) firstBad = i; bad++; }
            }
            say("  %-30s %9zu values  mismatches %zu  %s\n", t.n, vin.size(), bad, bad ? "FAIL" : "EXACT");
            if (bad)
                say("      first at i=%zu in=%u mask=%u got=%u\n", firstBad, vin[firstBad], mk[firstBad], vout[firstBad]);
        }
        say("================================================================\n");
    }

    if (c.vtopk) {  // mm:958-994
        say("\n=============== STEP B: parallel top-K verification ===============\n");
        say("  serial TG=256 threads | parallel TG=256 threads (max 1024)\n");
        const uint32_t TT = std::min(T, MAXT);
        std::vector<uint32_t> hs((size_t)TT * NCOL), hp((size_t)TT * NCOL);
        uint32_t mismatch = 0, thetaMismatch = 0;
        for (uint32_t q = 0; q < NSTIM; ++q) {
            CHECK(axon_net_set_topk_parallel(E.n, 0));
            E.present(q, 1.0, TT, CLAMP, false, true, 9000 + q);
            std::copy(E.hist.begin(), E.hist.begin() + hs.size(), hs.begin());
            uint32_t thS = E.theta[0];
            CHECK(axon_net_set_topk_parallel(E.n, 1));
            E.present(q, 1.0, TT, CLAMP, false, true, 9000 + q);
            std::copy(E.hist.begin(), E.hist.begin() + hp.size(), hp.begin());
            for (size_t i = 0; i < hs.size(); ++i)
                if (hs[i] != hp[i]) mismatch++;
            if (thS != E.theta[0]) thetaMismatch++;
        }
        say("  winner-set words compared: %zu   MISMATCHES: %u  -> %s\n", (size_t)TT * NCOL * NSTIM, mismatch,
               mismatch ? "FAIL" : "BIT-IDENTICAL");
        say("  final theta compared: %u stimuli, mismatches %u\n", NSTIM, thetaMismatch);
        auto timeIt = [&](bool par) {
            CHECK(axon_net_set_topk_parallel(E.n, par ? 1 : 0));
            double best = 1e30;
            for (int r = 0; r < 5; ++r) best = std::min(best, E.present(0, 1.0, TT, CLAMP, false, true, 9999));
            return best;
        };
        double ts = timeIt(false), tp = timeIt(true);
        say("  full step: serial %.1f us | parallel %.1f us | saved %.1f us (%.1f%% of step)\n", ts * 1000.0,
               tp * 1000.0, (ts - tp) * 1000.0, 100.0 * (ts - tp) / ts);
        CHECK(axon_net_set_topk_parallel(E.n, c.topkpar ? 1 : 0));
        say("==================================================================\n");
    }

    const double d0 = E.wstat();
    if (!quiet)
        say("initial functional synapse density: %.3f%%  (fan-in %.0f functional)\n\n", d0, d0 / 100.0 * FANIN);

    // settle homeostasis without learning (mm:1001); only the last one's counters are printed
    for (uint32_t i = 0; i < 10; ++i) E.present(i % NSTIM, 1.0, T, T, false, false, 999, 2u, i == 9);
    if (!quiet)
        say("settled: theta=%u  activeCols/step=%.0f  activeNeurons/step=%.0f\n\n", E.theta[0],
               (double)E.stats[0] / T, (double)E.stats[1] / T);

    // epochs (mm:1008-1024).  Presentations are enqueued back to back; the
    // counters are read at the end of each epoch that prints a line.
    double msAcc = 0;
    int msN = 0;
    if (skipEpochs) {
        // the step time comes from learning presentations on this untrained
        // network, which the trained one then replaces
        const uint32_t np = std::min(NSTIM, 12u);
        CHECK(axon_sync(E.n));
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t p = 0; p < np; ++p) E.present(p, 1.0, T, trainClamp, true, false, 1000 + p, 2u, false);
        CHECK(axon_sync(E.n));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        axon_net_destroy(E.n);
        E.n = nullptr;
        if (trained) {
            E.n = trained;
            E.owned = false;
        } else {
            CHECK(axon_net_load(loadPath, &E.n));
        }
        axon_config lc{};
        CHECK(axon_net_config(E.n, &lc));
        const axon_config a = training_keys(lc), b = training_keys(c);
        if (memcmp(&a, &b, sizeof a) != 0)
            throw RunError{trained ? AXON_ERR_INVALID_ARG : AXON_ERR_BAD_SNAPSHOT,
                           trained ? std::string("the trained network was created with other training keys than this "
                                                 "run's")
                                   : std::string(loadPath) + " was saved with other training keys than this run's"};
        if (!quiet) {
            if (trained) say("  trained network given (epochs not rerun)\n");
            else say("  loaded trained network %s (epochs not rerun)\n", loadPath);
        }
        msAcc = ms / T;
        msN = (int)np;
    }
    const auto tTrain = std::chrono::steady_clock::now();
    for (uint32_t ep = 0; ep < (skipEpochs ? 0u : EPOCHS); ++ep) {
        const bool pr = !quiet && (ep % 5 == 4 || ep == 0);
        for (uint32_t p = 0; p < NSTIM; ++p) {
            E.present(p, 1.0, T, trainClamp, true, false, 1000 + p, 2u, pr && p + 1 == NSTIM);
            msN++;
        }
        if (pr) {
            double ltp = E.stats[8], ltd = E.stats[9];
            uint32_t tp = E.stats[10], td = E.stats[11];
            double wd = E.wstat();
            say("  ep %3u  wdens %7.3f%%  LTP %10.0f  LTD %10.0f  LTD/LTP %6.3f  learners %u/%u\n", ep + 1, wd,
                   ltp, ltd, ltp > 0 ? ltd / ltp : 0.0, tp, td);
        }
    }
    CHECK(axon_sync(E.n));
    if (msN && !skipEpochs)
        msAcc = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tTrain).count() / T;
    if (msN == 0) { msAcc = E.present(0, 1.0, T, T, false, true, 777); msN = 1; }
    const double msStep = msAcc / msN;
    if (savePath) {
        CHECK(axon_net_save(E.n, savePath));
        if (!quiet) say("  saved trained network %s\n", savePath);
    }

    {  // firing steps per neuron, one extra training presentation (mm:1028-1072)
        E.present(0, 1.0, T, trainClamp, true, false, 31337);
        std::vector<uint16_t> cnt(N, 0), cntC(N, 0), cntF(N, 0);
        for (uint32_t t = 0; t < T; ++t)
            for (uint32_t col = IN; col < NCOL; ++col) {
                uint32_t b = E.row(t)[col];
                while (b) {
                    uint32_t l = __builtin_ctz(b); b &= b - 1;
                    cnt[col * NPC + l]++;
                    if (t < trainClamp) cntC[col * NPC + l]++; else cntF[col * NPC + l]++;
                }
            }
        const uint32_t* Hlast = E.row(T - 1);
        std::vector<uint8_t> isMember(N, 0);
        for (uint32_t col = IN; col < NCOL; ++col) {
            uint32_t b = Hlast[col];
            while (b) { uint32_t l = __builtin_ctz(b); b &= b - 1; isMember[col * NPC + l] = 1; }
        }
        double nn = 0, sum = 0, sumC = 0, sumF = 0, nm = 0, sm = 0, smC = 0, smF = 0;
        std::vector<uint16_t> v, vm;
        for (uint32_t i = IN * NPC; i < N; ++i) {
            if (cnt[i]) { nn++; sum += cnt[i]; sumC += cntC[i]; sumF += cntF[i]; v.push_back(cnt[i]); }
            if (isMember[i]) { nm++; sm += cnt[i]; smC += cntC[i]; smF += cntF[i]; vm.push_back(cnt[i]); }
        }
        std::sort(v.begin(), v.end());
        std::sort(vm.begin(), vm.end());
        say("\n================= FIRING STEPS PER NEURON (training) =================\n");
        say("  T=%u  trainClamp=%u (clamped) + %u (free)\n", T, trainClamp, T - trainClamp);
        say("  neurons firing at least once: %.0f\n", nn);
        say("  mean firing steps  : %.2f   (clamped %.2f + free %.2f)\n", nn ? sum / nn : 0.0, nn ? sumC / nn : 0.0,
               nn ? sumF / nn : 0.0);
        if (!v.empty())
            say("  p50 %u | p90 %u | max %u   | as fraction of T: %.3f\n", v[v.size() / 2],
                   v[(size_t)(0.9 * (v.size() - 1))], v.back(), nn ? sum / nn / T : 0.0);
        say("  --- restricted to FINAL-STATE ASSEMBLY MEMBERS (the allocated ones) ---\n");
        say("  members: %.0f | mean firing steps %.2f (clamped %.2f + free %.2f) | fraction of T %.3f\n", nm,
               nm ? sm / nm : 0.0, nm ? smC / nm : 0.0, nm ? smF / nm : 0.0, nm ? sm / nm / T : 0.0);
        if (!vm.empty())
            say("  p50 %u | p90 %u | max %u\n", vm[vm.size() / 2], vm[(size_t)(0.9 * (vm.size() - 1))], vm.back());
        say("  => bias gain per presentation = biasUp * spikes(members) = %.0f\n", nm ? biasUp * sm / nm : 0.0);
        say("======================================================================\n");
    }

    // references (mm:1080-1097)
    std::vector<Vec> refClampCols(NSTIM), refFreeCols(NSTIM), refClampBal(NSTIM), refFreeBal(NSTIM);
    std::vector<Vec> asmSpkHost(NSTIM, Vec(NCOL, 0));
    for (uint32_t p = 0; p < NSTIM; ++p) {
        E.present(p, 1.0, T, T, false, true, 2000 + p);
        refClampCols[p] = E.assemblyOf();
        refClampBal[p] = E.spk;
        asmSpkHost[p] = E.spk;
        for (uint32_t col = 0; col < IN; ++col) asmSpkHost[p][col] = 0u;
        E.present(p, 1.0, T, CLAMP, false, true, 2100 + p);
        refFreeCols[p] = E.assemblyOf();
        refFreeBal[p] = E.spk;
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

    if (c.diag) {  // mm:1107-1151
        Vec mp(N, 0), ap((size_t)NSTIM * NCOL, 0);
        for (uint32_t q = 0; q < NSTIM; ++q) std::copy(asmSpkHost[q].begin(), asmSpkHost[q].end(), ap.begin() + (size_t)q * NCOL);
        for (uint32_t col = IN; col < NCOL; ++col)
            for (uint32_t l = 0; l < NPC; ++l) {
                uint32_t m = 0;
                for (uint32_t q = 0; q < std::min(NSTIM, 32u); ++q)
                    if ((asmSpkHost[q][col] >> l) & 1u) m |= (1u << q);
                mp[col * NPC + l] = m;
            }
        uint32_t nMember = 0;
        for (uint32_t i = 0; i < N; ++i)
            if (mp[i]) nMember++;
        const char* selName[2] = {"RECURRENT branches", "AFFERENT branches"};
        say("\n======================= STEP 0 DIAGNOSTIC =======================\n");
        say("4-bit synaptic counter distribution, %u stimuli, %u assembly-member neurons\n", NSTIM, nMember);
        for (uint32_t sel = 0; sel < 2; ++sel) {
            uint32_t h[48];
            CHECK(axon_wdiag(E.n, mp.data(), ap.data(), std::min(NSTIM, 32u), sel, h));
            const char* cat[3] = {"within-assembly", "between-assembly", "all"};
            say("\n--- %s ---\n", selName[sel]);
            say("  %-17s %10s  %6s %6s %6s | %s\n", "pair type", "count", "mean", "%at 0", "%at 15", "histogram v=0..15");
            for (uint32_t cc = 0; cc < 3; ++cc) {
                double tot = 0, sum = 0;
                for (uint32_t v = 0; v < 16; ++v) { tot += h[cc * 16 + v]; sum += (double)v * h[cc * 16 + v]; }
                say("  %-17s %10.0f  %6.2f %6.1f %6.1f | ", cat[cc], tot, tot ? sum / tot : 0.0,
                       tot ? 100.0 * h[cc * 16 + 0] / tot : 0.0, tot ? 100.0 * h[cc * 16 + 15] / tot : 0.0);
                for (uint32_t v = 0; v < 16; ++v) say("%s%.1f", v ? "," : "", tot ? 100.0 * h[cc * 16 + v] / tot : 0.0);
                say("\n");
            }
        }
        say("================================================================\n");
    }

    if (!quiet) {  // mm:1153-1160
        say("\n  %.4f ms/timestep -> %.0f steps/s (%.1fx biological real time), wdens %.3f%%\n", msStep,
               1000.0 / msStep, 1.0 / msStep, E.wstat());
        say("  assembly size %.1f cols (%.2f%% of area) | cross-talk CLAMPED %.3f  FREE %.3f  (chance %.3f)\n\n", sz,
               100.0 * sz / (NCOL - IN), xtalk, xtalkF, sz / (NCOL - IN));
        say("--- pattern completion: cue %u steps, RELEASED %u.  BOTH reference definitions ---\n", CLAMP, T - CLAMP);
        say("  cue  |  CLAMPED ref: self  other  margin  |  FREE ref: self  other  margin\n");
    }
    if (c.profile)  // mm:1162-1340 used Metal timestamp counters (and mutated the weights)
        say("\n[profile] per-kernel GPU timing on MI355X comes from rocprofv3 --kernel-trace; no in-process profile\n");

    Vec ind(N, 0);
    {  // STEP A1 (mm:1342-1409)
        CHECK(axon_indeg(E.n, ind.data()));
        Vec v;
        v.reserve(N);
        for (uint32_t col = IN; col < NCOL; ++col)
            for (uint32_t l = 0; l < NPC; ++l) v.push_back(ind[col * NPC + l]);
        Vec sv = v;
        std::sort(sv.begin(), sv.end());
        auto pct = [&](double q) { return sv[(size_t)(q * (sv.size() - 1))]; };
        double mean = 0;
        for (uint32_t x : v) mean += x;
        mean /= v.size();
        double sd = 0;
        for (uint32_t x : v) sd += (x - mean) * (x - mean);
        sd = std::sqrt(sd / v.size());
        say("\n=============================== STEP A1 ===============================\n");
        say("Functional in-degree over the %u RECURRENT branches (popcount of MSB plane)\n", NSLOT - inBranches);
        say("  n=%zu neurons   mean %.1f   sd %.1f   cv %.3f\n", v.size(), mean, sd, sd / mean);
        say("  min %u | p10 %u | p25 %u | p50 %u | p75 %u | p90 %u | p99 %u | p99.9 %u | max %u\n", sv.front(),
               pct(0.10), pct(0.25), pct(0.50), pct(0.75), pct(0.90), pct(0.99), pct(0.999), sv.back());
        say("  tail ratio p99/p50 = %.2f   max/p50 = %.2f\n", (double)pct(0.99) / pct(0.50),
               (double)sv.back() / pct(0.50));
        Vec bias(N, 0);
        CHECK(axon_read_buffer(E.n, AXON_BUF_BIAS, bias.data(), (size_t)N * 4));
        for (int which = 0; which < 2; ++which) {
            const auto& src = which ? refClampBal : refFreeBal;
            std::vector<Vec> ns(NSTIM);
            for (uint32_t q = 0; q < NSTIM; ++q) ns[q] = E.nset(src[q].data());
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
                for (uint32_t l = 0; l < NPC; ++l) {
                    uint32_t id = col * NPC + l;
                    if (cnt[id] >= NSTIM - 1) { nAll++; dAll += ind[id]; bAll += bias[id]; }
                    if (cnt[id] == 0) { nNone++; dNone += ind[id]; bNone += bias[id]; }
                    if (cnt[id] > 0) { nAny++; dAny += ind[id]; bAny += bias[id]; }
                }
            say("\n  --- winner identity, %s endpoint ---\n", which ? "CLAMPED" : "FREE-RUN");
            say("    mean pairwise Jaccard %.3f (max %.3f) over %d pairs, mean set size %.0f\n", jm / jn, jmax, jn,
                   (double)ns[0].size());
            say("    neurons winning in >=%u/%u stimuli: %6u  in-degree %6.1f  intrinsic-bias %8.1f\n", NSTIM - 1,
                   NSTIM, nAll, nAll ? dAll / nAll : 0.0, nAll ? bAll / nAll : 0.0);
            say("    neurons winning in any stimulus  : %6u  in-degree %6.1f  intrinsic-bias %8.1f\n", nAny,
                   nAny ? dAny / nAny : 0.0, nAny ? bAny / nAny : 0.0);
            say("    neurons never winning            : %6u  in-degree %6.1f  intrinsic-bias %8.1f\n", nNone,
                   nNone ? dNone / nNone : 0.0, nNone ? bNone / nNone : 0.0);
            say("    HUB RATIO (winners-in-all / never) = %.3f\n",
                   (nNone && dNone) ? (dAll / std::max(1u, nAll)) / (dNone / nNone) : 0.0);
        }
        say("======================================================================\n");
    }

    std::vector<Vec> refNBCols(NSTIM);
    {  // STEP A3 + A4 (mm:1411-1511)
        const double npool = (double)(NCOL - IN) * NPC;
        std::vector<Vec> cns(NSTIM);
        for (uint32_t q = 0; q < NSTIM; ++q) cns[q] = E.nset(refClampBal[q].data());
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
                std::set_intersection(cns[i].begin(), cns[i].end(), cns[j].begin(), cns[j].end(), std::back_inserter(ii));
                totInt += ii.size();
            }
        Vec uni;
        for (auto& a : cns) uni.insert(uni.end(), a.begin(), a.end());
        std::sort(uni.begin(), uni.end());
        size_t uniqN = std::unique(uni.begin(), uni.end()) - uni.begin();
        double sumSz = msz * NSTIM;
        say("\n=============================== STEP A3 ===============================\n");
        say("Chance baseline for %u assemblies of %.0f neurons from a pool of %.0f\n", NSTIM, msz, npool);
        say("  E[|A n B|] = m^2/n            = %.2f neurons\n", eInt);
        say("  E[Jaccard] = E/(2m-E)         = %.5f      <-- CHANCE\n", eJac);
        say("  measured mean pairwise Jaccard = %.5f   (max %.5f)\n", jm / jn, jmax);
        say("  ratio measured/chance          = %.3f\n", eJac > 0 ? (jm / jn) / eJac : 0.0);
        say("  total pairwise overlap: measured %.0f neurons vs %.0f expected at chance (%.1fx less)\n", totInt,
               eInt * jn, totInt > 0 ? eInt * jn / totInt : 1e9);
        say("  sum of assembly sizes %.0f | union %zu | excess %.0f  -> %s\n", sumSz, uniqN, sumSz - (double)uniqN,
               (sumSz - (double)uniqN) < 0.02 * sumSz ? "NEAR-EXACT PARTITION" : "overlapping");

        for (uint32_t q = 0; q < NSTIM; ++q) {
            E.present(q, 1.0, T, T, false, true, 2200 + q, 0u);
            refNBCols[q] = E.assemblyOf();
        }
        double preserve = 0;
        for (uint32_t q = 0; q < NSTIM; ++q) preserve += overlap(refNBCols[q], refClampCols[q]);
        preserve /= NSTIM;
        double xNB = 0;
        int xn2 = 0;
        for (uint32_t i = 0; i < NSTIM; ++i)
            for (uint32_t j = i + 1; j < NSTIM; ++j) { xNB += overlap(refNBCols[i], refNBCols[j]); xn2++; }
        xNB /= xn2;
        say("\n=============================== STEP A4 ===============================\n");
        say("Zero the intrinsic bias at TEST time only. Trained weights unchanged.\n");
        say("  clamped assembly preserved when bias is zeroed: overlap %.3f\n", preserve);
        say("  cross-talk of clamped assemblies, bias OFF     : %.3f  (bias ON: %.3f)\n", xNB, xtalk);
        say("\n  release trajectory, cue 75%%, bias OFF at test (vs bias ON baseline)\n");
        say("  step  phase   ov(clampRef,biasON)  ov(clampRef,biasOFF)  ov(noBiasRef,biasOFF)\n");
        std::vector<double> a1(T, 0), a2(T, 0), a3(T, 0);
        for (uint32_t q = 0; q < NSTIM; ++q) {
            E.present(q, 0.75, T, CLAMP, false, true, 4300 + q, 1u);
            for (uint32_t t = 0; t < T; ++t) a1[t] += overlap(E.cols_of(E.row(t)), refClampCols[q]);
            E.present(q, 0.75, T, CLAMP, false, true, 4300 + q, 0u);
            for (uint32_t t = 0; t < T; ++t) {
                auto a = E.cols_of(E.row(t));
                a2[t] += overlap(a, refClampCols[q]);
                a3[t] += overlap(a, refNBCols[q]);
            }
        }
        for (uint32_t t = 0; t < T; ++t)
            say("  %4u  %-5s %19.3f %21.3f %22.3f\n", t, t < CLAMP ? "cue" : "FREE", a1[t] / NSTIM, a2[t] / NSTIM,
                   a3[t] / NSTIM);
        say("\n  pattern completion with bias OFF at test  (keep=0 is the no-cue control)\n");
        say("  cue  |  vs CLAMPED ref(biasON): self  other  margin | vs noBias ref: self  other  margin\n");
        for (double keep : {1.0, 0.75, 0.5, 0.35, 0.25, 0.15, 0.10, 0.05, 0.02, 0.0}) {
            double s1 = 0, o1 = 0, s2 = 0, o2 = 0;
            for (uint32_t q = 0; q < NSTIM; ++q) {
                E.present(q, keep, T, CLAMP, false, true, 4400 + q, 0u);
                auto a = E.assemblyOf();
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
            say("  %3.0f%% |%25.3f %6.3f %+7.3f |%18.3f %6.3f %+7.3f\n", keep * 100, s1 / NSTIM, o1 / NSTIM,
                   s1 / NSTIM - o1 / NSTIM, s2 / NSTIM, o2 / NSTIM, s2 / NSTIM - o2 / NSTIM);
        }
        say("======================================================================\n");
    }

    {  // A4b + A4c (mm:1513-1599)
        Vec unt(N, 0);
        CHECK(axon_indeg(E.n, ind.data()));
        CHECK(axon_untouched(E.n, unt.data()));
        const double RECSYN = (double)(NSLOT - inBranches) * PRESYN;
        const double cues[4] = {1.0, 0.75, 0.35, 0.15};
        say("\n=============================== A4b ===============================\n");
        say("Is the recalled ~25%% a FIXED subset? (readout gate on: bias off at test)\n");
        std::vector<std::vector<Vec>> rec(NSTIM, std::vector<Vec>(4));
        for (uint32_t q = 0; q < NSTIM; ++q)
            for (int ci = 0; ci < 4; ++ci) {
                E.present(q, cues[ci], T, CLAMP, false, true, 5000 + q * 16 + ci);
                rec[q][ci] = E.nset(E.spk.data());
            }
        say("  pairwise Jaccard of recalled sets ACROSS cue levels, same stimulus:\n    ");
        double jall = 0;
        int jn = 0;
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                double m = 0;
                for (uint32_t q = 0; q < NSTIM; ++q) m += jacc(rec[q][a], rec[q][b]);
                m /= NSTIM;
                say("%.0f%%/%.0f%%=%.3f  ", cues[a] * 100, cues[b] * 100, m);
                jall += m; jn++;
            }
        say("\n    mean across cue pairs = %.3f\n", jall / jn);
        double coreSz = 0, periSz = 0, asmSz = 0, dCore = 0, dPeri = 0, uCore = 0, uPeri = 0;
        size_t nCore = 0, nPeri = 0;
        for (uint32_t q = 0; q < NSTIM; ++q) {
            Vec core = rec[q][0];
            for (int ci = 1; ci < 4; ++ci) {
                Vec t;
                std::set_intersection(core.begin(), core.end(), rec[q][ci].begin(), rec[q][ci].end(), std::back_inserter(t));
                core = t;
            }
            Vec A = E.nset(refClampBal[q].data()), coreInA, peri;
            std::set_intersection(core.begin(), core.end(), A.begin(), A.end(), std::back_inserter(coreInA));
            std::set_difference(A.begin(), A.end(), core.begin(), core.end(), std::back_inserter(peri));
            coreSz += coreInA.size(); periSz += peri.size(); asmSz += A.size();
            for (uint32_t x : coreInA) { dCore += ind[x]; uCore += 100.0 * unt[x] / RECSYN; nCore++; }
            for (uint32_t x : peri) { dPeri += ind[x]; uPeri += 100.0 * unt[x] / RECSYN; nPeri++; }
        }
        say("\n  assembly %.0f neurons | CORE (recalled at all 4 cues) %.0f (%.1f%%) | periphery %.0f\n",
               asmSz / NSTIM, coreSz / NSTIM, 100.0 * coreSz / asmSz, periSz / NSTIM);
        say("  %-12s %10s %14s %22s\n", "group", "n", "in-degree", "%% synapses still at 5/6");
        say("  %-12s %10zu %14.1f %22.1f\n", "CORE", nCore, nCore ? dCore / nCore : 0.0, nCore ? uCore / nCore : 0.0);
        say("  %-12s %10zu %14.1f %22.1f\n", "periphery", nPeri, nPeri ? dPeri / nPeri : 0.0, nPeri ? uPeri / nPeri : 0.0);
        double allD = 0, allU = 0;
        size_t nAll = 0;
        for (uint32_t col = IN; col < NCOL; ++col)
            for (uint32_t l = 0; l < NPC; ++l) {
                allD += ind[col * NPC + l]; allU += 100.0 * unt[col * NPC + l] / RECSYN; nAll++;
            }
        say("  %-12s %10zu %14.1f %22.1f\n", "all neurons", nAll, allD / nAll, allU / nAll);
        Vec bs(N, 0);
        CHECK(axon_read_buffer(E.n, AXON_BUF_BIAS, bs.data(), (size_t)N * 4));
        Vec bv;
        bv.reserve(N);
        for (uint32_t col = IN; col < NCOL; ++col)
            for (uint32_t l = 0; l < NPC; ++l) bv.push_back(bs[col * NPC + l]);
        std::sort(bv.begin(), bv.end());
        double bm = 0;
        for (uint32_t x : bv) bm += x;
        bm /= bv.size();
        say("\n=============================== A4c ===============================\n");
        say("Intrinsic bias growth after %u epochs (unbounded, non-leaky, learning-time only)\n", EPOCHS);
        say("  mean %.0f | p50 %u | p99 %u | max %u | UINT32_MAX %u | headroom %.3g x\n", bm, bv[bv.size() / 2],
               bv[(size_t)(0.99 * (bv.size() - 1))], bv.back(), 4294967295u, bv.back() ? 4294967295.0 / bv.back() : 0.0);
        say("  growth rate ~= %.1f per epoch (max/epochs) -> epochs to saturate ~ %.3g\n",
               EPOCHS ? (double)bv.back() / EPOCHS : 0.0,
               (bv.back() && EPOCHS) ? 4294967295.0 / ((double)bv.back() / EPOCHS) : 0.0);
        say("==================================================================\n");
    }

    Vec mult(N, 0);  // bMult mirror (zero until unionchar/split fill it)
    const uint32_t T35 = std::min(35u, MAXT);
    if (c.unionchar) {  // mm:1601-1652
        std::fill(mult.begin(), mult.end(), 0u);
        for (uint32_t q = 0; q < NSTIM; ++q)
            for (uint32_t x : E.nset(refClampBal[q].data())) mult[x]++;
        CHECK(axon_set_mult(E.n, mult.data()));
        const uint32_t NQ = std::min(NSTIM, 12u);
        double nIn = 0, nOut = 0, fI = 0, pI = 0, fO = 0, pO = 0, fA = 0, pA = 0, mOut = 0, mAll = 0, nAll = 0;
        Vec postMask(N, 0);
        Vec allRec;
        for (uint32_t col = IN; col < NCOL; ++col)
            for (uint32_t l = 0; l < NPC; ++l) allRec.push_back(col * NPC + l);
        for (uint32_t q = 0; q < NQ; ++q) {
            CHECK(axon_set_ref_ball(E.n, refClampBal[q].data()));
            E.present(q, 1.0, T35, CLAMP, false, true, 9500 + q);
            Vec uni;
            for (uint32_t t = CLAMP; t < T35; ++t)
                for (uint32_t x : E.nset(E.row(t))) uni.push_back(x);
            std::sort(uni.begin(), uni.end());
            uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
            Vec ref = E.nset(refClampBal[q].data()), in_, out_;
            std::set_intersection(uni.begin(), uni.end(), ref.begin(), ref.end(), std::back_inserter(in_));
            std::set_difference(uni.begin(), uni.end(), ref.begin(), ref.end(), std::back_inserter(out_));
            nIn += in_.size(); nOut += out_.size();
            // k_within's presynaptic ballot is bRefBall = refClampBal[q] (mm:1617)
            auto runQ = [&](const Vec& post, double& f, double& pp) {
                std::fill(postMask.begin(), postMask.end(), 0u);
                for (uint32_t x : post) postMask[x] = 1u;
                uint32_t wv[16];
                CHECK(axon_within(E.n, refClampBal[q].data(), postMask.data(), wv));
                f += wv[0]; pp += wv[1];
            };
            runQ(in_, fI, pI); runQ(out_, fO, pO); runQ(allRec, fA, pA);
            for (uint32_t x : out_) mOut += mult[x];
            for (uint32_t x : allRec) mAll += mult[x];
            nAll += allRec.size();
        }
        say("\nUNIONCHAR %.0f %.0f %.4f %.4f %.4f %.4f %.4f\n", nIn / NQ, nOut / NQ, 100.0 * fI / pI,
               100.0 * fO / pO, 100.0 * fA / pA, nOut > 0 ? mOut / nOut : 0.0, nAll > 0 ? mAll / nAll : 0.0);
    }

    if (c.split) {  // mm:1654-1692
        std::fill(mult.begin(), mult.end(), 0u);
        for (uint32_t q = 0; q < NSTIM; ++q)
            for (uint32_t x : E.nset(refClampBal[q].data())) mult[x]++;
        CHECK(axon_set_mult(E.n, mult.data()));
        const uint32_t NQ = std::min(NSTIM, 12u);
        double fS = 0, pS = 0, fE = 0, pE = 0, nS = 0, nE = 0;
        std::vector<double> dS(64, 0), cS(64, 0), dE(64, 0), cE(64, 0);
        for (uint32_t q = 0; q < NQ; ++q) {
            CHECK(axon_set_ref_ball(E.n, refClampBal[q].data()));
            uint32_t wv[16];
            CHECK(axon_within(E.n, refClampBal[q].data(), nullptr, wv));
            fS += wv[4]; pS += wv[5]; fE += wv[6]; pE += wv[7]; nS += wv[8]; nE += wv[9];
            CHECK(axon_clear_drive_log(E.n));
            E.present(q, 1.0, T35, CLAMP, false, true, 8800 + q);
            for (uint32_t t = 0; t < T35; ++t) {
                dS[t] += E.dlog[t * 4 + 0]; cS[t] += E.dlog[t * 4 + 1];
                dE[t] += E.dlog[t * 4 + 2]; cE[t] += E.dlog[t * 4 + 3];
            }
        }
        say("\n===== SHARED vs EXCLUSIVE assembly members (NSTIM=%u) =====\n", NSTIM);
        say("  shared fraction of assembly : %.3f  (%.0f shared / %.0f exclusive neurons)\n",
               (nS + nE) > 0 ? nS / (nS + nE) : 0.0, nS / NQ, nE / NQ);
        say("  within-assembly density  SHARED    %.4f%%  (%.0f/%.0f)\n", 100.0 * fS / pS, fS, pS);
        say("  within-assembly density  EXCLUSIVE %.4f%%  (%.0f/%.0f)\n", 100.0 * fE / pE, fE, pE);
        say("  free-phase drive  SHARED   :");
        for (uint32_t t = CLAMP; t < T35; ++t) say(" %.0f", cS[t] ? dS[t] / cS[t] : 0.0);
        say("\n  free-phase drive  EXCLUSIVE:");
        for (uint32_t t = CLAMP; t < T35; ++t) say(" %.0f", cE[t] ? dE[t] / cE[t] : 0.0);
        say("\n==========================================================\n");
    }

    if (c.deep) {  // mm:1694-1730
        const uint32_t NQ = std::min(NSTIM, 12u);
        const uint32_t TT = std::min(std::min(T + 20u, 40u), MAXT);
        double fIn = 0, pIn = 0, fOut = 0, pOut = 0;
        std::vector<double> drvSum(40, 0), drvCnt(40, 0);
        for (uint32_t q = 0; q < NQ; ++q) {
            CHECK(axon_set_ref_ball(E.n, refClampBal[q].data()));
            uint32_t wv[16];
            CHECK(axon_within(E.n, refClampBal[q].data(), nullptr, wv));
            fIn += wv[0]; pIn += wv[1]; fOut += wv[2]; pOut += wv[3];
            CHECK(axon_clear_drive_log(E.n));
            E.present(q, 1.0, TT, CLAMP, false, true, 8600 + q);
            for (uint32_t t = 0; t < TT; ++t) {
                drvSum[t] += E.dlog[t * 4 + 0] + E.dlog[t * 4 + 2];
                drvCnt[t] += E.dlog[t * 4 + 1] + E.dlog[t * 4 + 3];
            }
        }
        say("\n========= WITHIN-ASSEMBLY RECURRENT SCAFFOLD (NSTIM=%u) =========\n", NSTIM);
        say("  within-assembly : %.0f functional of %.0f possible = %.4f%%\n", fIn, pIn, 100.0 * fIn / pIn);
        say("  between         : %.0f functional of %.0f possible = %.4f%%\n", fOut, pOut, 100.0 * fOut / pOut);
        say("  enrichment within/between = %.2fx\n", (fOut > 0 && pIn > 0) ? (fIn / pIn) / (fOut / pOut) : 0.0);
        say("  free-phase drive to assembly members, per step:\n   ");
        for (uint32_t t = 0; t < TT; ++t)
            if (t < CLAMP + 12) say(" t%u:%.0f", t, drvCnt[t] ? drvSum[t] / drvCnt[t] : 0.0);
        say("\n================================================================\n");
    }

    if (c.deep) {  // mm:1732-1801
        std::vector<uint16_t> mlt(N, 0);
        for (uint32_t q = 0; q < NSTIM; ++q)
            for (uint32_t x : E.nset(refClampBal[q].data())) mlt[x]++;
        double popSum = 0;
        uint32_t popN = 0, maxMult = 0;
        for (uint32_t i = IN * NPC; i < N; ++i)
            if (mlt[i]) { popSum += mlt[i]; popN++; maxMult = std::max<uint32_t>(maxMult, mlt[i]); }
        double popMean = popN ? popSum / popN : 0.0;
        const uint32_t NQ = std::min(NSTIM, 30u);
        double mIn = 0, mOut = 0, nIn = 0, nOut = 0, sumBest = 0, sumMeanOther = 0, sumSelf = 0;
        for (uint32_t q = 0; q < NQ; ++q) {
            E.present(q, 1.0, T, CLAMP, false, true, 8100 + q);
            Vec recN = E.nset(E.spk.data()), refN = E.nset(refClampBal[q].data()), in_, out_;
            std::set_intersection(recN.begin(), recN.end(), refN.begin(), refN.end(), std::back_inserter(in_));
            std::set_difference(recN.begin(), recN.end(), refN.begin(), refN.end(), std::back_inserter(out_));
            for (uint32_t x : in_) { mIn += mlt[x]; nIn++; }
            for (uint32_t x : out_) { mOut += mlt[x]; nOut++; }
            Vec recCols = E.assemblyOf();
            sumSelf += overlap(recCols, refClampCols[q]);
            double best = 0, tot = 0;
            uint32_t cnt = 0;
            for (uint32_t r = 0; r < NSTIM; ++r)
                if (r != q) {
                    double o = overlap(recCols, refClampCols[r]);
                    best = std::max(best, o); tot += o; cnt++;
                }
            sumBest += best;
            sumMeanOther += cnt ? tot / cnt : 0.0;
        }
        say("\n============ RETRIEVED-POPULATION MULTIPLICITY (NSTIM=%u) ============\n", NSTIM);
        say("  population: %u neurons used, mean multiplicity %.2f, max %u\n", popN, popMean, maxMult);
        say("  recalled AND in reference    : n=%.0f  mean multiplicity %.2f\n", nIn, nIn ? mIn / nIn : 0.0);
        say("  recalled but NOT in reference: n=%.0f  mean multiplicity %.2f  (%.2fx population)\n", nOut,
               nOut ? mOut / nOut : 0.0, (nOut && popMean) ? (mOut / nOut) / popMean : 0.0);
        say("  self %.3f | best-other %.3f | MEAN-over-others %.3f  -> %s\n", sumSelf / NQ, sumBest / NQ,
               sumMeanOther / NQ, (sumMeanOther / NQ > 0.5 * sumBest / NQ) ? "UNIFORM contamination" : "TAIL-dominated");
        auto jac2 = [&](const Vec& a, const Vec& b) {
            Vec i, u;
            std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(i));
            std::set_union(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(u));
            return u.empty() ? 0.0 : (double)i.size() / u.size();
        };
        std::vector<Vec> ns(NSTIM);
        for (uint32_t q = 0; q < NSTIM; ++q) ns[q] = E.nset(refClampBal[q].data());
        std::vector<double> J, IO, IX;
        for (uint32_t i = 0; i < NSTIM; ++i)
            for (uint32_t j = i + 1; j < NSTIM; ++j) {
                J.push_back(jac2(ns[i], ns[j]));
                Vec a = stimCols[i], b = stimCols[j];
                std::sort(a.begin(), a.end());
                std::sort(b.begin(), b.end());
                Vec ii;
                std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(ii));
                IO.push_back((double)ii.size());
                IX.push_back((double)std::max(i, j));
            }
        auto pear = [&](std::vector<double>& x, std::vector<double>& y) {
            double mx = 0, my = 0;
            for (size_t i = 0; i < x.size(); ++i) { mx += x[i]; my += y[i]; }
            mx /= x.size(); my /= y.size();
            double sxy = 0, sxx = 0, syy = 0;
            for (size_t i = 0; i < x.size(); ++i) {
                double dx = x[i] - mx, dy = y[i] - my;
                sxy += dx * dy; sxx += dx * dx; syy += dy * dy;
            }
            return (sxx > 0 && syy > 0) ? sxy / std::sqrt(sxx * syy) : 0.0;
        };
        if (!J.empty()) {
            std::vector<double> Js = J;
            std::sort(Js.begin(), Js.end());
            say("  heavy tail: %zu pairs | mean J %.5f | p99 %.5f | max %.5f\n", J.size(),
                   std::accumulate(J.begin(), J.end(), 0.0) / J.size(), Js[(size_t)(0.99 * (Js.size() - 1))], Js.back());
            say("  corr(assemblyJaccard, INPUT overlap)      r = %+.3f\n", pear(J, IO));
            say("  corr(assemblyJaccard, allocation index)   r = %+.3f\n", pear(J, IX));
        }
        say("======================================================================\n");
    }

    if (c.wave) {  // mm:1803-1858
        const uint32_t TW = std::min(std::min(c.wavet, 60u), MAXT);
        const uint32_t NQ = std::min(NSTIM, 12u);
        std::vector<double> ovR(TW, 0), nCol(TW, 0), nNeu(TW, 0), jPrev(TW, 0), fracA(TW, 0), jPrev2(TW, 0);
        double unionSz = 0, unionE = 0, unionO = 0, unionX = 0;
        for (uint32_t q = 0; q < NQ; ++q) {
            E.present(q, 1.0, TW, CLAMP, false, true, 7000 + q);
            Vec refN = E.nset(refClampBal[q].data());
            Vec prev, prev2, uni, uniE, uniO;
            for (uint32_t t = 0; t < TW; ++t) {
                const uint32_t* row = E.row(t);
                Vec cur = E.nset(row);
                uint32_t nc = 0;
                for (uint32_t col = IN; col < NCOL; ++col)
                    if (row[col]) nc++;
                nCol[t] += nc; nNeu[t] += cur.size();
                Vec ii;
                std::set_intersection(cur.begin(), cur.end(), refN.begin(), refN.end(), std::back_inserter(ii));
                ovR[t] += (cur.empty() || refN.empty()) ? 0.0 : (double)ii.size() / std::sqrt((double)cur.size() * refN.size());
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
        say("\nPARITY %.0f %.4f %.0f %.4f %.0f %.4f %.4f\n", unionSz / NQ, unionSz / NQ / 1536.0, unionE / NQ,
               unionE / NQ / 1536.0, unionO / NQ, unionO / NQ / 1536.0, unionX / NQ);
        say("\n================= WAVE: %u free steps after a %u-step cue =================\n", TW - CLAMP, CLAMP);
        say("  step  phase   cols  neurons   ov(clampRef)  fracOfAssembly  J(t,t-1)   J(t,t-2)\n");
        for (uint32_t t = 0; t < TW; ++t)
            say("  %4u  %-5s %6.0f %8.0f %13.3f %15.3f %10.3f %10.3f\n", t, t < CLAMP ? "cue" : "FREE", nCol[t] / NQ,
                   nNeu[t] / NQ, ovR[t] / NQ, fracA[t] / NQ, jPrev[t] / NQ, jPrev2[t] / NQ);
        say("==========================================================================\n");
    }

    {  // free-run diagnostic (mm:1860-1887)
        say("\n============ FREE-RUN DIAGNOSTIC (cue %u steps, then released) ============\n", CLAMP);
        say("  cue=75%%, averaged over %u stimuli.  'margin' = k-th minus (k+1)-th drive\n", NSTIM);
        say("  step  phase  cols  drive_k  margin  %%marg=0   ov(CLAMPED ref)  ov(FREE ref)\n");
        std::vector<double> ovt(T, 0), mg(T, 0), dk(T, 0), tie(T, 0), zer(T, 0), nc(T, 0), ovf(T, 0);
        for (uint32_t q = 0; q < NSTIM; ++q) {
            E.present(q, 0.75, T, CLAMP, false, true, 4000 + q);
            const uint32_t* S = E.stats.data();
            for (uint32_t t = 0; t < T; ++t) {
                auto a = E.cols_of(E.row(t));
                ovt[t] += overlap(a, refClampCols[q]);
                ovf[t] += overlap(a, refFreeCols[q]);
                const uint32_t* sb = S + 64 + t * 8;
                double nn = sb[2];
                nc[t] += nn;
                if (nn > 0) {
                    mg[t] += sb[0] / nn; dk[t] += sb[1] / nn; tie[t] += 100.0 * sb[3] / nn; zer[t] += 100.0 * sb[4] / nn;
                }
            }
        }
        for (uint32_t t = 0; t < T; ++t)
            say("  %4u  %-5s %5.0f %8.1f %7.1f %8.1f   %14.3f  %12.3f\n", t, t < CLAMP ? "cue" : "FREE", nc[t] / NSTIM,
                   dk[t] / NSTIM, mg[t] / NSTIM, tie[t] / NSTIM, ovt[t] / NSTIM, ovf[t] / NSTIM);
        say("==========================================================================\n");
    }

    double m50 = 0;  // mm:1889-1906
    for (double keep : {1.0, 0.75, 0.5, 0.35, 0.25}) {
        double self = 0, other = 0, selfF = 0, otherF = 0;
        for (uint32_t p = 0; p < NSTIM; ++p) {
            E.present(p, keep, T, CLAMP, false, true, 3000 + p);
            auto a = E.assemblyOf();
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
        if (keep == 0.5) m50 = self - other;
        if (!quiet)
            say("  %3.0f%% |  %16.3f %6.3f %+7.3f  |  %11.3f %6.3f %+7.3f\n", keep * 100, self, other, self - other,
                   selfF, otherF, selfF - otherF);
    }
    if (quiet)
        say("tgt=%-4u k=%-3u inbr=%-3u stimc=%-4u bias=%-4u nand=%u | %7.4f ms %6.0f st/s asm=%-5.0f xtalk=%.3f "
               "margin50=%+.3f wd=%.2f\n",
               targetC, K, inBranches, c.stimcols, biasUp, c.nand, msStep, 1000.0 / msStep, sz, xtalk, m50, E.wstat());
    else
        say("\n=== DONE ===\n");
    fflush(g_out);
}

extern "C" axon_status axon_run(int argc, const char* const* argv, axon_net* trained, const char* report_path) {
    if (argc < 1 || !argv) return axon::set_error(AXON_ERR_INVALID_ARG, "argv must hold at least the program name");
    FILE* out = stdout;
    if (report_path && !(out = fopen(report_path, "w")))
        return axon::set_error(AXON_ERR_IO, std::string("cannot create ") + report_path);
    FILE* prev = g_out;
    g_out = out;
    axon_status st = AXON_OK;
    std::string msg;
    try {
        run(argc, argv, trained);
    } catch (const RunError& e) {
        st = e.s;
        msg = e.msg;
    } catch (const std::bad_alloc&) {
        st = AXON_ERR_OUT_OF_MEMORY;
        msg = "host memory exhausted";
    } catch (...) {
        st = AXON_ERR_DEVICE;
        msg = "unexpected exception";
    }
    g_out = prev;
    if (out != stdout && fclose(out) != 0 && st == AXON_OK) return axon::set_error(AXON_ERR_IO, "writing the report failed");
    if (out == stdout) fflush(stdout);
    return st == AXON_OK ? AXON_OK : axon::set_error(st, msg);
}
