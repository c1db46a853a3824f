// C-ABI host layer of the AXON engine (libaxon.so): configuration, device
// memory, the presentation sequencer and the readouts.  See axon_api.h.
//
// A presentation (present(), mm:815-888) is T steps of drive -> top-K ->
// emit(+bias+plasticity).  The per-step parameters depend only on (T, clamp
// steps, learn, k, facil, useBias), so the whole presentation is captured once
// per such key in a hipGraph and replayed; the stimulus is the only per-call
// input and is staged into device memory before the replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <cerrno>
#include <random>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/axon_api.h"
#include "engine.h"
#include "host_rng.h"

using axon::NetDev;
using axon::Params;

namespace {

thread_local std::string g_err;

axon_status fail(axon_status s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
axon_status fail(axon_status s, const char* fmt, ...) {
    char b[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(b, sizeof b, fmt, ap);
    va_end(ap);
    g_err = b;
    return s;
}

}  // namespace

namespace axon {
// Sets the calling thread's axon_last_error() message for the other source
// files of the library (report.cpp).
axon_status set_error(axon_status s, const std::string& msg) {
    g_err = msg;
    return s;
}
}  // namespace axon

namespace {

// Entry points that size host memory from their arguments or a file run
// through this, so an exception never crosses the C ABI: host memory
// exhaustion becomes AXON_ERR_OUT_OF_MEMORY.
template <class F>
axon_status guarded(const char* what, F&& f) noexcept {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(AXON_ERR_OUT_OF_MEMORY, "%s: host memory exhausted", what);
    } catch (const std::exception& e) {
        return fail(AXON_ERR_DEVICE, "%s: %s", what, e.what());
    } catch (...) {
        return fail(AXON_ERR_DEVICE, "%s: unexpected exception", what);
    }
}

#define HIPCHECK(x)                                                                                  \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(AXON_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_));     \
    } while (0)

constexpr uint32_t NSTAGE = 8;  // pinned stimulus staging slots

// One device, checked once: the code objects in this library are gfx950 only.
axon_status select_device(int* dev) {
    static int cached = -2;
    static std::string why;
    if (cached == -2) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            cached = -1;
            why = "no HIP device visible";
        } else {
            int d = 0;
            if (const char* e = getenv("AXON_DEVICE")) d = atoi(e);
            hipDeviceProp_t p;
            if (d >= n || hipGetDeviceProperties(&p, d) != hipSuccess) {
                cached = -1;
                why = "cannot query HIP device";
            } else if (strncmp(p.gcnArchName, "gfx950", 6) != 0) {
                cached = -1;
                why = std::string("device is ") + p.gcnArchName + ", this build targets gfx950 (MI355X)";
            } else {
                cached = d;
            }
        }
    }
    if (cached < 0) return fail(AXON_ERR_DEVICE, "%s", why.c_str());
    if (hipSetDevice(cached) != hipSuccess) return fail(AXON_ERR_DEVICE, "hipSetDevice failed");
    *dev = cached;
    return AXON_OK;
}

uint32_t argu(int argc, const char* const* argv, const char* key, uint32_t dflt) {  // mm:660-667
    const size_t kl = strlen(key);
    uint32_t v = dflt;
    for (int i = 1; i < argc; ++i)
        if (!strncmp(argv[i], key, kl) && argv[i][kl] == '=') v = (uint32_t)atoi(argv[i] + kl + 1);
    return v;
}

struct GraphKey {
    uint32_t T, clamp, learn, k, facil, useBias, stats, step0;
    uint32_t row;  // 1 + the pattern whose table row step 1 reads, or 0 (the stimulus is copied)
    bool operator<(const GraphKey& o) const {
        return std::tie(T, clamp, learn, k, facil, useBias, stats, step0, row) <
               std::tie(o.T, o.clamp, o.learn, o.k, o.facil, o.useBias, o.stats, o.step0, o.row);
    }
};

}  // namespace

struct axon_net {
    axon_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    NetDev d{};
    uint32_t NCOL = 0, N = 0, MAXT = 0, NSTATS = 0, NDLOG = 0;
    size_t PLANE_U4 = 0;
    std::vector<std::vector<uint32_t>> stimCols;
    Params P{};
    uint32_t K = 0, kTest = 0, facil = 0, contrast = 0, ltdEvery = 1;
    uint32_t* h_stim[NSTAGE] = {};
    hipEvent_t stim_ev[NSTAGE] = {};
    bool stim_ev_live[NSTAGE] = {};
    uint32_t stage = 0;
    uint32_t* d_tab = nullptr;    // the patterns' stimuli at keep 1.0, NCOL words each (built at first use)
    uint32_t* d_small = nullptr;  // 64 words of readout counters
    uint32_t* d_ball = nullptr;   // NCOL words
    uint32_t* d_mask = nullptr;   // N words
    std::map<GraphKey, hipGraphExec_t> graphs;
    bool use_graphs = true;
    // per-kernel event timing (axon_net_set_profiling)
    bool profiling = false;
    std::vector<hipEvent_t> prof_ev;  // 2 per kernel per step
    double prof_ms[4] = {0, 0, 0, 0};
    uint64_t prof_n[4] = {0, 0, 0, 0};
    std::vector<void*> allocs;
};

// ================================================================== config
extern "C" uint32_t axon_abi_version(void) { return AXON_ABI_VERSION; }

extern "C" const char* axon_status_str(axon_status s) {
    switch (s) {
        case AXON_OK: return "ok";
        case AXON_ERR_INVALID_CONFIG: return "invalid config";
        case AXON_ERR_OUT_OF_MEMORY: return "out of device memory";
        case AXON_ERR_BUDGET_EXCEEDED: return "HBM budget exceeded";
        case AXON_ERR_DEVICE: return "device error";
        case AXON_ERR_INVALID_ARG: return "invalid argument";
        case AXON_ERR_CAPACITY: return "capacity exceeded";
        case AXON_ERR_NOT_READY: return "not ready";
        case AXON_ERR_IO: return "file error";
        case AXON_ERR_BAD_SNAPSHOT: return "bad snapshot";
    }
    return "unknown status";
}

extern "C" const char* axon_last_error(void) { return g_err.c_str(); }

extern "C" void axon_config_defaults(axon_config* c) {  // mm:688-728 (built-in defaults, not Config C)
    memset(c, 0, sizeof(*c));
    c->nand = 4; c->target = 256; c->k = 6; c->dthr = 2; c->boost = 4; c->mode = 1; c->epochs = 30;
    c->T = 10; c->clampsteps = 5; c->stim = 12; c->quiet = 0; c->seed = 0; c->split = 0; c->unionchar = 0;
    c->biasup = 0; c->biasdown = 1; c->biasshift = 4; c->biasdownevery = 1; c->inbr = 8; c->stimcols = 128;
    c->incols = 512; c->ncol = 4096; c->diag = 0; c->profile = 0; c->topkpar = 1; c->vtopk = 0;
    c->contrast = 0; c->bandlo = 5; c->bandhi = 11; c->margingate = 1; c->hetero = 0; c->vbs = 0;
    c->ltdrep = 1; c->ltdevery = 1; c->ktest = 0; c->facil = 0; c->deep = 0; c->wave = 0; c->wavet = 35;
    c->trainclamp = c->contrast ? c->clampsteps : c->T;
    c->max_t = 64;
}

extern "C" axon_status axon_config_parse(int argc, const char* const* argv, axon_config* c) {
    if (!c || (argc > 0 && !argv)) return fail(AXON_ERR_INVALID_ARG, "null argument");
    axon_config_defaults(c);
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
    c->trainclamp = argu(argc, argv, "trainclamp", c->contrast ? c->clampsteps : c->T);  // mm:731
    if (c->biasup == 1u && c->ncol > c->incols) {                                        // mm:732-735
        const double pr = (double)c->target * c->k / (double)(c->ncol - c->incols) / axon::NPC;
        c->biasup = (uint32_t)std::max(2.0, (1.0 - pr) / pr * c->biasdown);
    }
    c->max_t = argu(argc, argv, "maxt", 64);
    return AXON_OK;
}

// Rejects the configurations on which the reference has undefined behaviour
// (e.g. incols >= ncol divides by zero in k_init_conn, mm:116) instead of
// running them.
extern "C" axon_status axon_config_validate(const axon_config* c) {
    if (!c) return fail(AXON_ERR_INVALID_ARG, "null config");
    auto bad = [](const char* m) { return fail(AXON_ERR_INVALID_CONFIG, "%s", m); };
    if (c->ncol == 0 || c->ncol % axon::COLS_PER_GRP) return bad("ncol must be a positive multiple of 8");
    if (c->ncol > (1u << 20)) return bad("ncol must be <= 1048576");
    if (c->incols == 0 || c->incols % axon::COLS_PER_GRP) return bad("incols must be a positive multiple of 8");
    if (c->incols >= c->ncol) return bad("incols must be < ncol");
    if (c->stimcols > c->incols) return bad("stimcols must be <= incols");
    if (c->inbr > axon::NSLOT) return bad("inbr must be <= 32");
    if (c->biasshift >= 32) return bad("biasshift must be < 32");
    if (c->contrast && c->ltdevery == 0) return bad("ltdevery must be >= 1");
    if (c->max_t == 0 || c->max_t > 4096) return bad("max_t must be in 1..4096");
    if (c->stim == 0) return bad("stim must be >= 1");
    return AXON_OK;
}

// ================================================================== network
static axon_status dev_alloc(axon_net* n, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(AXON_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    n->allocs.push_back(*p);
    return AXON_OK;
}

extern "C" void axon_net_destroy(axon_net* n) {
    if (!n) return;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (auto& kv : n->graphs) (void)hipGraphExecDestroy(kv.second);
    for (uint32_t i = 0; i < NSTAGE; ++i) {
        if (n->stim_ev[i]) (void)hipEventDestroy(n->stim_ev[i]);
        if (n->h_stim[i]) (void)hipHostFree(n->h_stim[i]);
    }
    for (hipEvent_t e : n->prof_ev) (void)hipEventDestroy(e);
    for (void* p : n->allocs) (void)hipFree(p);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

static axon_status net_create(const axon_config* c, axon_net** out) {
    if (!c || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    axon_status st = axon_config_validate(c);
    if (st != AXON_OK) return st;
    int dev = 0;
    if ((st = select_device(&dev)) != AXON_OK) return st;
    axon_net* n = new axon_net();
    n->cfg = *c;
    n->device = dev;
    n->NCOL = c->ncol;
    n->N = c->ncol * axon::NPC;
    n->MAXT = c->max_t;
    n->NSTATS = std::max<uint32_t>(1024u, 64u + 8u * n->MAXT);
    n->NDLOG = std::max<uint32_t>(256u, 4u * n->MAXT);
    n->PLANE_U4 = (size_t)n->NCOL * axon::NSLOT * (axon::TILE_WORDS / 4u);
    if (const char* e = getenv("AXON_NO_GRAPH")) n->use_graphs = atoi(e) == 0;

    auto cleanup = [&](axon_status s) { axon_net_destroy(n); return s; };
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(AXON_ERR_DEVICE, "hipStreamCreate failed"));
    NetDev& d = n->d;
    void* p = nullptr;
    uint32_t** planes[4] = {&d.W0, &d.W1, &d.W2, &d.W3};
    for (int i = 0; i < 4; ++i) {
        if ((st = dev_alloc(n, &p, n->PLANE_U4 * 16)) != AXON_OK) return cleanup(st);
        *planes[i] = (uint32_t*)p;
    }
    struct { uint32_t** ptr; size_t words; } bufs[] = {
        {&d.conn, (size_t)n->NCOL * axon::NSLOT}, {&d.bias, n->N}, {&d.spk[0], n->NCOL}, {&d.spk[1], n->NCOL},
        {&d.acc, n->N}, {&d.bbits, n->N}, {&d.drive, n->NCOL}, {&d.theta, 16}, {&d.marg, n->NCOL},
        {&d.stim, n->NCOL}, {&d.stats, n->NSTATS}, {&d.hist, (size_t)n->MAXT * n->NCOL}, {&d.driveLog, n->NDLOG},
        {&d.refBall, n->NCOL}, {&d.mult, n->N}, {&n->d_small, 64}, {&n->d_ball, n->NCOL}, {&n->d_mask, n->N},
        {&d.cntProp, (size_t)n->NCOL * 2}, {&d.cntLearn, (size_t)n->NCOL * 2},
        {(uint32_t**)&d.firedW, (size_t)n->NCOL * axon::NSLOT / 4}, {(uint32_t**)&d.tileBand, (size_t)n->NCOL * axon::NSLOT / 4},
        {(uint32_t**)&d.colNZ, (n->NCOL + 16) / 4},
        {&d.learnList, (size_t)2 * axon::NLISTS * axon::list_cols(n->NCOL)},
        {&d.learnCnt, 2 * axon::NLISTS * 32}, {&d.cntLast, (size_t)n->NCOL * 2},
    };
    for (auto& b : bufs) {
        if ((st = dev_alloc(n, &p, b.words * 4)) != AXON_OK) return cleanup(st);
        *b.ptr = (uint32_t*)p;
        // Metal zero-fills new buffers; bias, refBall, mult and driveLog rely on it (mm:758-763)
        if (hipMemsetAsync(p, 0, b.words * 4, n->stream) != hipSuccess)
            return cleanup(fail(AXON_ERR_DEVICE, "hipMemsetAsync failed"));
    }
    d.ncol = n->NCOL; d.incols = c->incols; d.inbr = c->inbr; d.topkpar = c->topkpar;
    d.nstats = n->NSTATS; d.ndlog = n->NDLOG; d.maxt = n->MAXT;
    for (uint32_t i = 0; i < NSTAGE; ++i) {
        if (hipHostMalloc((void**)&n->h_stim[i], (size_t)n->NCOL * 4, hipHostMallocDefault) != hipSuccess)
            return cleanup(fail(AXON_ERR_OUT_OF_MEMORY, "hipHostMalloc failed"));
        if (hipEventCreateWithFlags(&n->stim_ev[i], hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(AXON_ERR_DEVICE, "hipEventCreate failed"));
    }
    if (axon::launch_init(d, c->nand, c->seed, n->stream) != hipSuccess)
        return cleanup(fail(AXON_ERR_DEVICE, "init kernels failed to launch"));
    if (hipStreamSynchronize(n->stream) != hipSuccess)
        return cleanup(fail(AXON_ERR_DEVICE, "init kernels failed"));

    // stimulus columns, mm:793-799
    std::mt19937 rng(7u + c->seed * 104729u);
    n->stimCols.resize(c->stim);
    std::vector<uint32_t> all(c->incols);
    for (uint32_t q = 0; q < c->stim; ++q) {
        for (uint32_t i = 0; i < c->incols; ++i) all[i] = i;
        axon::libcxx_shuffle(all.data(), c->incols, rng);
        n->stimCols[q].assign(all.begin(), all.begin() + c->stimcols);
    }
    // Params fixed for the network, mm:810-812
    Params& P = n->P;
    P = Params{};
    P.k = c->k; P.dthr = c->dthr; P.boost = c->boost; P.mode = c->mode; P.inCols = c->incols;
    P.targetCols = c->target; P.biasUp = c->biasup; P.biasDown = c->biasdown; P.biasShift = c->biasshift;
    P.bandLo = c->bandlo; P.bandHi = c->contrast ? c->bandhi : 0u; P.marginGate = c->margingate;
    P.hetero = c->hetero; P.ltdRep = c->ltdrep; P.biasDownEvery = c->biasdownevery; P.histOn = 1u;
    n->K = c->k; n->kTest = c->ktest; n->facil = c->facil; n->contrast = c->contrast;
    n->ltdEvery = c->ltdevery ? c->ltdevery : 1u;
    // a neuron's accumulator is at most 32 branches x (256 presynaptic bits +
    // boost) in mode 1 (x 256 in mode 0, x 1 in mode 2) plus the facilitation
    const uint64_t perBranch = P.mode == 0u ? 256u : P.mode == 1u ? 256ull + P.boost : 1u;
    d.acc16 = (uint64_t)axon::NSLOT * perBranch + n->facil <= 0xFFFFu ? 1u : 0u;
    *out = n;
    return AXON_OK;
}

// ================================================================== present
// The phase and the other per-step fields of step t (mm:837-884).
static Params step_params(const axon_net* n, Params P, uint32_t t, uint32_t clampSteps, bool learn) {
    P.step = t;
    P.learn = learn ? 1u : 0u;
    P.freeze = learn ? 0u : 1u;
    P.clamp = (t < clampSteps) ? 1u : 0u;
    P.depress = (learn && (t % 2u == 1u)) ? 1u : 0u;
    P.phase = (!learn || !n->contrast) ? 0u
              : (t < clampSteps)      ? 1u
              : (((t - clampSteps) % n->ltdEvery == 0u) ? 2u : 3u);
    return P;
}

// The launches of one presentation.  Without side statistics, step 0 of a
// presentation of two or more steps is one launch (k_emit_first: the spikes
// start cleared, so that step's propagation yields zeros and its plasticity
// has no voted tile; the buffers the skipped launches would have left --
// accumulators, drives, NMDA bits of step 0 -- are overwritten by step 1),
// and the plasticity of step t runs inside step t + 1's propagation, the
// last step's in a launch of its own.  With side statistics every step is
// clear/propagation/emit/plasticity.  ev (optional): 8 timing events per
// step -- propagation, first-step emit, emit, plasticity -- and plast[t]
// records which steps launched the plasticity kernel.
// step0: step 0 runs inside step 1's propagation (a first-launch
// presentation, clamped at step 0 to a stimulus inside the input columns),
// which reads the stimulus from stim0 (d.stim when null, or a row of the
// pattern table) and keeps it in d.stim.
static hipError_t enqueue_steps(axon_net* n, Params P0, uint32_t T, uint32_t clampSteps, bool learn, bool stats,
                                bool step0, const uint32_t* stim0 = nullptr, hipEvent_t* ev = nullptr,
                                uint8_t* plast = nullptr) {
    const bool first = !stats && T >= 2u;
    step0 = step0 && first && clampSteps >= 1u;
    hipError_t e = first ? hipSuccess : axon::launch_clear(n->d, n->stream);
    Params PL{};
    bool pending = false;  // the previous step's plasticity waits for this step's propagation
    auto E = [&](uint32_t t, int k) -> hipEvent_t { return ev ? ev[(size_t)8 * t + k] : nullptr; };
    for (uint32_t t = 0; t < T && e == hipSuccess; ++t) {
        const Params P = step_params(n, P0, t, clampSteps, learn);
        if (first && t == 0u) {
            if (!step0) e = axon::launch_emit(n->d, P, t, learn, stats, n->stream, E(t, 2), E(t, 3), true);
            continue;
        }
        const Params S0 = step_params(n, P0, 0u, clampSteps, learn);
        e = axon::launch_drive(n->d, P, t, n->stream, E(t, 0), E(t, 1), pending ? &PL : nullptr,
                               step0 && t == 1u ? &S0 : nullptr, stim0);
        pending = false;
        if (e == hipSuccess && !stats && learn && P.phase != 3u && t + 1u == T) {
            // the last step's plasticity in the same launch as its emit
            e = axon::launch_emit_learn(n->d, P, t, n->stream, E(t, 6), E(t, 7));
            if (plast) plast[t] = 2u;
            continue;
        }
        if (e == hipSuccess) e = axon::launch_emit(n->d, P, t, learn, stats, n->stream, E(t, 4), E(t, 5));
        if (e != hipSuccess || !learn || P.phase == 3u) continue;
        if (stats) {
            e = axon::launch_learn(n->d, P, t, n->stream, E(t, 6), E(t, 7));
            if (plast) plast[t] = 1u;
        } else {
            PL = P;
            pending = true;
        }
    }
    return e;  // nothing pending: the last step's plasticity ran with its emit, or the step did not learn
}

// The launches of enqueue_steps, each carrying start / end timing events, so
// the kernels' device durations can be summed after the presentation.
static hipError_t enqueue_steps_profiled(axon_net* n, Params P, uint32_t T, uint32_t clampSteps, bool learn,
                                         bool stats, bool step0, const uint32_t* stim0) {
    const size_t need = (size_t)8 * std::max<uint32_t>(T, 1);
    while (n->prof_ev.size() < need) {
        hipEvent_t e;
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
        n->prof_ev.push_back(e);
    }
    std::vector<uint8_t> plast(T, 0);  // steps with a plasticity launch: 1 its own, 2 with the emit
    hipError_t e = enqueue_steps(n, P, T, clampSteps, learn, stats, step0, stim0, n->prof_ev.data(), plast.data());
    if (e != hipSuccess) return e;
    if ((e = hipStreamSynchronize(n->stream)) != hipSuccess) return e;
    const bool first = !stats && T >= 2u;  // as enqueue_steps
    step0 = step0 && first && clampSteps >= 1u;
    for (uint32_t t = step0 ? 1u : 0u; t < T; ++t) {
        const bool one = first && t == 0u;  // the first-step launch alone
        for (int k = 0; k < 4; ++k) {
            if (k == 3 ? !plast[t] : (one != (k == 1)) || (k == 2 && plast[t] == 2u)) continue;
            float ms = 0.f;
            if ((e = hipEventElapsedTime(&ms, n->prof_ev[(size_t)8 * t + 2 * k], n->prof_ev[(size_t)8 * t + 2 * k + 1])) != hipSuccess)
                return e;
            n->prof_ms[k] += ms;
            n->prof_n[k] += 1;
        }
    }
    return hipSuccess;
}

extern "C" axon_status axon_present(axon_net* n, const axon_present_args* a, axon_present_out* o) {
    if (!n || !a) return fail(AXON_ERR_INVALID_ARG, "null argument");
    if (a->T > n->MAXT)
        return fail(AXON_ERR_CAPACITY, "T=%u exceeds the history capacity max_t=%u", a->T, n->MAXT);
    if (!a->stim && a->pattern >= n->stimCols.size())
        return fail(AXON_ERR_INVALID_ARG, "pattern %u out of range (stim=%zu)", a->pattern, n->stimCols.size());
    if (a->use_bias > 2u) return fail(AXON_ERR_INVALID_ARG, "use_bias must be 0, 1 or 2");
    const bool async = (a->flags & AXON_PRESENT_ASYNC) != 0u;
    if (async && o && (o->spikes || o->hist || o->stats || o->drive_log || o->theta))
        return fail(AXON_ERR_INVALID_ARG, "AXON_PRESENT_ASYNC takes no outputs");
    HIPCHECK(hipSetDevice(n->device));
    const bool learn = a->learn != 0u;
    Params P = n->P;  // mm:818-822
    P.k = learn ? n->K : (n->kTest != 0u ? n->kTest : n->K);
    P.facil = learn ? 0u : n->facil;
    P.useBias = (a->use_bias == 2u) ? (learn ? 1u : 0u) : a->use_bias;

    // Stimulus: loadStim (mm:800-804).  Step 0 of a first-launch presentation
    // clamped at step 0 runs inside step 1's propagation when the stimulus lies
    // inside the input columns (then it is step 0's spike vector).  A stored
    // pattern at keep 1.0 (every column of the pattern, whatever the seed:
    // canonical01 < 1) is then read from the pattern table, with no copy; any
    // other stimulus goes through a pinned staging slot into d.stim.
    bool step0 = async && a->T >= 2u && a->clamp_steps >= 1u;
    const uint32_t* stim0 = nullptr;
    uint32_t row = 0u;
    if (step0 && !a->stim && a->keep >= 1.0f) {
        if (!n->d_tab) {
            const uint32_t np = (uint32_t)n->stimCols.size();
            std::vector<uint32_t> tab((size_t)np * n->NCOL, 0u);
            for (uint32_t q = 0; q < np; ++q)
                for (uint32_t c : n->stimCols[q]) tab[(size_t)q * n->NCOL + c] = 0xFFFFFFFFu;
            void* p = nullptr;
            const axon_status st = dev_alloc(n, &p, tab.size() * 4);
            if (st != AXON_OK) return st;
            n->d_tab = static_cast<uint32_t*>(p);
            HIPCHECK(hipMemcpyAsync(p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, n->stream));
            HIPCHECK(hipStreamSynchronize(n->stream));  // tab goes away on return
        }
        row = a->pattern + 1u;
        stim0 = n->d_tab + (size_t)a->pattern * n->NCOL;
    } else {
        const uint32_t slot = n->stage;  // the presentation's until the device has passed the copy (stim_ev)
        n->stage = (n->stage + 1u) % NSTAGE;
        if (n->stim_ev_live[slot]) HIPCHECK(hipEventSynchronize(n->stim_ev[slot]));
        uint32_t* hs = n->h_stim[slot];
        if (a->stim) {
            memcpy(hs, a->stim, (size_t)n->NCOL * 4);
        } else {
            memset(hs, 0, (size_t)n->NCOL * 4);
            std::mt19937 r(a->seed);
            for (uint32_t c : n->stimCols[a->pattern])
                if (axon::canonical01(r) <= a->keep) hs[c] = 0xFFFFFFFFu;
        }
        for (uint32_t c = n->cfg.incols; step0 && c < n->NCOL; ++c) step0 = hs[c] == 0u;
        HIPCHECK(hipMemcpyAsync(n->d.stim, hs, (size_t)n->NCOL * 4, hipMemcpyHostToDevice, n->stream));
        HIPCHECK(hipEventRecord(n->stim_ev[slot], n->stream));
        n->stim_ev_live[slot] = true;
    }

    const auto t0 = std::chrono::steady_clock::now();
    if (n->profiling) {
        // waits for the device even when async (the event times are read back)
        hipError_t e = enqueue_steps_profiled(n, P, a->T, a->clamp_steps, learn, !async, step0, stim0);
        if (e != hipSuccess) return fail(AXON_ERR_DEVICE, "profiled launch: %s", hipGetErrorString(e));
    } else if (n->use_graphs && a->T > 0) {
        const GraphKey key{a->T, a->clamp_steps, learn ? 1u : 0u, P.k, P.facil, P.useBias, async ? 0u : 1u,
                           step0 ? 1u : 0u, row};
        auto it = n->graphs.find(key);
        if (it == n->graphs.end()) {
            hipGraph_t g = nullptr;
            HIPCHECK(hipStreamBeginCapture(n->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t e = enqueue_steps(n, P, a->T, a->clamp_steps, learn, !async, step0, stim0);
            hipError_t e2 = hipStreamEndCapture(n->stream, &g);
            if (e != hipSuccess) return fail(AXON_ERR_DEVICE, "capture launch: %s", hipGetErrorString(e));
            if (e2 != hipSuccess) return fail(AXON_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e2));
            hipGraphExec_t ex = nullptr;
            hipError_t e3 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e3 != hipSuccess) return fail(AXON_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e3));
            it = n->graphs.emplace(key, ex).first;
        }
        HIPCHECK(hipGraphLaunch(it->second, n->stream));
    } else {
        const hipError_t e = enqueue_steps(n, P, a->T, a->clamp_steps, learn, !async, step0, stim0);
        if (e != hipSuccess) return fail(AXON_ERR_DEVICE, "launch: %s", hipGetErrorString(e));
    }
    if (async) {
        if (o) o->ms_per_step = 0.0;
        return AXON_OK;
    }
    HIPCHECK(hipStreamSynchronize(n->stream));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (o) {
        o->ms_per_step = a->T ? ms / a->T : 0.0;
        const size_t NC = n->NCOL;
        if (o->spikes) HIPCHECK(hipMemcpy(o->spikes, n->d.spk[a->T & 1u], NC * 4, hipMemcpyDeviceToHost));
        if (o->hist && a->T) HIPCHECK(hipMemcpy(o->hist, n->d.hist, (size_t)a->T * NC * 4, hipMemcpyDeviceToHost));
        if (o->stats) HIPCHECK(hipMemcpy(o->stats, n->d.stats, (size_t)n->NSTATS * 4, hipMemcpyDeviceToHost));
        if (o->drive_log) HIPCHECK(hipMemcpy(o->drive_log, n->d.driveLog, (size_t)n->NDLOG * 4, hipMemcpyDeviceToHost));
        if (o->theta) HIPCHECK(hipMemcpy(o->theta, n->d.theta, 8, hipMemcpyDeviceToHost));
    }
    return AXON_OK;
}

extern "C" axon_status axon_sync(axon_net* n) {
    if (!n) return fail(AXON_ERR_INVALID_ARG, "null net");
    HIPCHECK(hipStreamSynchronize(n->stream));
    return AXON_OK;
}

// ================================================================== buffers
static axon_status buffer_ptr(const axon_net* n, axon_buffer w, void** p, size_t* words) {
    const NetDev& d = n->d;
    const size_t pw = n->PLANE_U4 * 4;
    switch (w) {
        case AXON_BUF_W0: *p = d.W0; *words = pw; break;
        case AXON_BUF_W1: *p = d.W1; *words = pw; break;
        case AXON_BUF_W2: *p = d.W2; *words = pw; break;
        case AXON_BUF_W3: *p = d.W3; *words = pw; break;
        case AXON_BUF_CONN: *p = d.conn; *words = (size_t)n->NCOL * axon::NSLOT; break;
        case AXON_BUF_SPK_A: *p = d.spk[0]; *words = n->NCOL; break;
        case AXON_BUF_SPK_B: *p = d.spk[1]; *words = n->NCOL; break;
        case AXON_BUF_DRIVE: *p = d.drive; *words = n->NCOL; break;
        case AXON_BUF_BBITS: *p = d.bbits; *words = n->N; break;
        case AXON_BUF_ACC: *p = d.acc; *words = n->N; break;
        case AXON_BUF_BIAS: *p = d.bias; *words = n->N; break;
        case AXON_BUF_MARG: *p = d.marg; *words = n->NCOL; break;
        case AXON_BUF_REFBALL: *p = d.refBall; *words = n->NCOL; break;
        case AXON_BUF_DRIVELOG: *p = d.driveLog; *words = n->NDLOG; break;
        case AXON_BUF_MULT: *p = d.mult; *words = n->N; break;
        case AXON_BUF_STIM: *p = d.stim; *words = n->NCOL; break;
        case AXON_BUF_STATS: *p = d.stats; *words = n->NSTATS; break;
        case AXON_BUF_HIST: *p = d.hist; *words = (size_t)n->MAXT * n->NCOL; break;
        case AXON_BUF_THETA: *p = d.theta; *words = 16; break;
        default: return fail(AXON_ERR_INVALID_ARG, "unknown buffer %d", (int)w);
    }
    return AXON_OK;
}

extern "C" axon_status axon_buffer_words(const axon_net* n, axon_buffer w, size_t* words) {
    if (!n || !words) return fail(AXON_ERR_INVALID_ARG, "null argument");
    void* p;
    return buffer_ptr(n, w, &p, words);
}

// The weight planes are exchanged with the host in the reference's layout
// (per tile: [half][neuron][4 words], mm:170-172); on the device each tile is
// word-major ([word][neuron]).  Tiles are 1 KiB and sit at the same offsets
// in both layouts, so a conversion permutes within whole tiles.
static bool is_plane(axon_buffer w) {
    return w == AXON_BUF_W0 || w == AXON_BUF_W1 || w == AXON_BUF_W2 || w == AXON_BUF_W3;
}
// Host-to-device copies and fills of a network's buffers go on its stream:
// the stream is non-blocking, so work on the null stream is not ordered
// before its kernels, and a fill issued there can land after a later kernel.
static hipError_t put(axon_net* n, void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, n->stream);
    return e == hipSuccess ? hipStreamSynchronize(n->stream) : e;  // src may go away on return
}
static hipError_t zero(axon_net* n, void* dst, size_t bytes) { return hipMemsetAsync(dst, 0, bytes, n->stream); }

static inline size_t ref_to_dev_word(size_t i) {  // index within the plane
    const size_t tile = i >> 8, r = i & 255u;
    const size_t h = r >> 7, nn = (r >> 2) & 31u, k = r & 3u;
    return (tile << 8) | ((h * 4u + k) << 5) | nn;
}

extern "C" axon_status axon_read_buffer(axon_net* n, axon_buffer w, void* dst, size_t bytes) {
    if (!n || !dst) return fail(AXON_ERR_INVALID_ARG, "null argument");
    void* p; size_t words;
    axon_status st = buffer_ptr(n, w, &p, &words);
    if (st != AXON_OK) return st;
    if (bytes > words * 4) return fail(AXON_ERR_INVALID_ARG, "read of %zu bytes exceeds buffer (%zu)", bytes, words * 4);
    HIPCHECK(hipSetDevice(n->device));
    HIPCHECK(hipStreamSynchronize(n->stream));
    if (w == AXON_BUF_ACC && n->d.acc16) {  // 16-bit accumulators, widened
        std::vector<uint16_t> h(bytes / 4);
        HIPCHECK(hipMemcpy(h.data(), p, h.size() * 2, hipMemcpyDeviceToHost));
        uint32_t* o = static_cast<uint32_t*>(dst);
        for (size_t i = 0; i < h.size(); ++i) o[i] = h[i];
        return AXON_OK;
    }
    if (!is_plane(w)) {
        HIPCHECK(hipMemcpy(dst, p, bytes, hipMemcpyDeviceToHost));
        return AXON_OK;
    }
    const size_t tw = axon::TILE_WORDS, ntw = (bytes / 4 + tw - 1) / tw * tw;  // whole tiles covering the read
    std::vector<uint32_t> dv(ntw), rv(ntw);
    HIPCHECK(hipMemcpy(dv.data(), p, ntw * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ntw; ++i) rv[i] = dv[ref_to_dev_word(i)];
    memcpy(dst, rv.data(), bytes);
    return AXON_OK;
}

extern "C" axon_status axon_write_buffer(axon_net* n, axon_buffer w, const void* src, size_t bytes) {
    if (!n || !src) return fail(AXON_ERR_INVALID_ARG, "null argument");
    void* p; size_t words;
    axon_status st = buffer_ptr(n, w, &p, &words);
    if (st != AXON_OK) return st;
    if (bytes > words * 4) return fail(AXON_ERR_INVALID_ARG, "write of %zu bytes exceeds buffer (%zu)", bytes, words * 4);
    HIPCHECK(hipSetDevice(n->device));
    HIPCHECK(hipStreamSynchronize(n->stream));
    if (w == AXON_BUF_ACC && n->d.acc16) {  // 16-bit accumulators: values above 65535 do not fit
        const uint32_t* v = static_cast<const uint32_t*>(src);
        std::vector<uint16_t> h(bytes / 4);
        for (size_t i = 0; i < h.size(); ++i) {
            if (v[i] > 0xFFFFu) return fail(AXON_ERR_INVALID_ARG, "accumulator %zu = %u does not fit 16 bits", i, v[i]);
            h[i] = (uint16_t)v[i];
        }
        HIPCHECK(put(n, p, h.data(), h.size() * 2));
        return AXON_OK;
    }
    if (!is_plane(w)) {
        HIPCHECK(put(n, p, src, bytes));
        return AXON_OK;
    }
    // read-modify-write of the covered tiles, then forget which tiles were known to lie in the band
    const size_t tw = axon::TILE_WORDS, ntw = (bytes / 4 + tw - 1) / tw * tw;
    std::vector<uint32_t> dv(ntw), rv(ntw);
    HIPCHECK(hipMemcpy(dv.data(), p, ntw * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ntw; ++i) rv[i] = dv[ref_to_dev_word(i)];
    memcpy(rv.data(), src, bytes);
    for (size_t i = 0; i < ntw; ++i) dv[ref_to_dev_word(i)] = rv[i];
    HIPCHECK(put(n, p, dv.data(), ntw * 4));
    HIPCHECK(zero(n, n->d.tileBand, (size_t)n->NCOL * axon::NSLOT));
    return AXON_OK;
}

extern "C" axon_status axon_set_ref_ball(axon_net* n, const uint32_t* ball) {
    return axon_write_buffer(n, AXON_BUF_REFBALL, ball, (size_t)(n ? n->NCOL : 0) * 4);
}
extern "C" axon_status axon_set_mult(axon_net* n, const uint32_t* mult) {
    return axon_write_buffer(n, AXON_BUF_MULT, mult, (size_t)(n ? n->N : 0) * 4);
}
extern "C" axon_status axon_clear_drive_log(axon_net* n) {
    if (!n) return fail(AXON_ERR_INVALID_ARG, "null net");
    HIPCHECK(hipSetDevice(n->device));
    HIPCHECK(hipMemsetAsync(n->d.driveLog, 0, (size_t)n->NDLOG * 4, n->stream));
    return AXON_OK;
}

extern "C" axon_status axon_net_set_topk_parallel(axon_net* n, uint32_t par) {
    if (!n) return fail(AXON_ERR_INVALID_ARG, "null net");
    HIPCHECK(hipSetDevice(n->device));
    HIPCHECK(hipStreamSynchronize(n->stream));
    n->d.topkpar = par ? 1u : 0u;
    for (auto& kv : n->graphs) (void)hipGraphExecDestroy(kv.second);  // kernel args captured the old mode
    n->graphs.clear();
    return AXON_OK;
}

// Sums the per-column counter pairs: out[0] += entries 2c, out[1] += 2c + 1.
static axon_status counter_pairs(axon_net* n, uint32_t* dev, uint64_t out[2], uint32_t reset) {
    std::vector<uint32_t> v((size_t)n->NCOL * 2);
    HIPCHECK(hipMemcpy(v.data(), dev, v.size() * 4, hipMemcpyDeviceToHost));
    out[0] = 0; out[1] = 0;
    for (size_t i = 0; i < v.size(); ++i) out[i & 1u] += v[i];
    if (reset) HIPCHECK(zero(n, dev, v.size() * 4));
    return AXON_OK;
}

// The per-column work counter pairs: propagation (cntProp) and plasticity,
// whose work at a presentation's last timestep (in the emit launch) counts
// apart (cntLast) from the rest (cntLearn).
static axon_status work_pairs(axon_net* n, uint64_t prop[2], uint64_t learn[2], uint64_t last[2], uint32_t reset) {
    HIPCHECK(hipSetDevice(n->device));
    HIPCHECK(hipStreamSynchronize(n->stream));
    axon_status st = counter_pairs(n, n->d.cntProp, prop, 0);
    if (st == AXON_OK) st = counter_pairs(n, n->d.cntLearn, learn, 0);
    if (st == AXON_OK) st = counter_pairs(n, n->d.cntLast, last, 0);
    if (st != AXON_OK) return st;
    if (reset) {
        HIPCHECK(zero(n, n->d.cntProp, (size_t)n->NCOL * 8));
        HIPCHECK(zero(n, n->d.cntLearn, (size_t)n->NCOL * 8));
        HIPCHECK(zero(n, n->d.cntLast, (size_t)n->NCOL * 8));
    }
    return AXON_OK;
}

extern "C" axon_status axon_work_counters(axon_net* n, uint64_t out[2], uint32_t reset) {
    if (!n || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    uint64_t p[2], l[2], z[2];
    const axon_status st = work_pairs(n, p, l, z, reset);
    if (st != AXON_OK) return st;
    out[0] = p[0]; out[1] = l[0] + z[0];
    return AXON_OK;
}

extern "C" axon_status axon_line_counters(axon_net* n, uint64_t out[2], uint32_t reset) {
    if (!n || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    uint64_t p[2], l[2], z[2];
    const axon_status st = work_pairs(n, p, l, z, reset);
    if (st != AXON_OK) return st;
    out[0] = p[1]; out[1] = l[1] + z[1];
    return AXON_OK;
}

extern "C" axon_status axon_last_step_counters(axon_net* n, uint64_t out[2]) {
    if (!n || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    uint64_t p[2], l[2];
    return work_pairs(n, p, l, out, 0);
}

extern "C" axon_status axon_net_set_profiling(axon_net* n, uint32_t enable) {
    if (!n) return fail(AXON_ERR_INVALID_ARG, "null net");
    n->profiling = enable != 0u;
    return AXON_OK;
}

extern "C" axon_status axon_kernel_times(axon_net* n, double ms[4], uint64_t launches[4], uint32_t reset) {
    if (!n || !ms || !launches) return fail(AXON_ERR_INVALID_ARG, "null argument");
    for (int k = 0; k < 4; ++k) {
        ms[k] = n->prof_ms[k];
        launches[k] = n->prof_n[k];
        if (reset) { n->prof_ms[k] = 0; n->prof_n[k] = 0; }
    }
    return AXON_OK;
}

extern "C" axon_status axon_stim_cols(const axon_net* n, uint32_t p, uint32_t* out) {
    if (!n || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    if (p >= n->stimCols.size()) return fail(AXON_ERR_INVALID_ARG, "pattern %u out of range", p);
    std::copy(n->stimCols[p].begin(), n->stimCols[p].end(), out);
    return AXON_OK;
}

// ================================================================== readouts
extern "C" axon_status axon_wstat(axon_net* n, uint32_t* out) {
    if (!n || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    HIPCHECK(hipSetDevice(n->device));
    HIPCHECK(hipMemsetAsync(n->d_small, 0, 4, n->stream));
    HIPCHECK(axon::launch_wstat(n->d, n->d_small, n->stream));
    HIPCHECK(hipMemcpyAsync(out, n->d_small, 4, hipMemcpyDeviceToHost, n->stream));
    HIPCHECK(hipStreamSynchronize(n->stream));
    return AXON_OK;
}

extern "C" axon_status axon_within(axon_net* n, const uint32_t* ball, const uint32_t* post_mask, uint32_t out[16]) {
    if (!n || !ball || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    HIPCHECK(hipSetDevice(n->device));
    HIPCHECK(hipMemcpyAsync(n->d_ball, ball, (size_t)n->NCOL * 4, hipMemcpyHostToDevice, n->stream));
    if (post_mask) HIPCHECK(hipMemcpyAsync(n->d_mask, post_mask, (size_t)n->N * 4, hipMemcpyHostToDevice, n->stream));
    HIPCHECK(hipMemsetAsync(n->d_small, 0, 64, n->stream));
    HIPCHECK(axon::launch_within(n->d, n->d_ball, n->d_mask, post_mask ? 1u : 0u, n->d_small, n->stream));
    uint32_t tmp[16];
    HIPCHECK(hipMemcpyAsync(tmp, n->d_small, 64, hipMemcpyDeviceToHost, n->stream));
    HIPCHECK(hipStreamSynchronize(n->stream));
    memcpy(out, tmp, sizeof tmp);  // the reference zeroes bWCnt before each call (mm:1701)
    return AXON_OK;
}

static axon_status per_neuron(axon_net* n, uint32_t* out, hipError_t (*fn)(const NetDev&, uint32_t*, hipStream_t)) {
    if (!n || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    HIPCHECK(hipSetDevice(n->device));
    HIPCHECK(fn(n->d, n->d_mask, n->stream));
    HIPCHECK(hipMemcpyAsync(out, n->d_mask, (size_t)n->N * 4, hipMemcpyDeviceToHost, n->stream));
    HIPCHECK(hipStreamSynchronize(n->stream));
    return AXON_OK;
}
extern "C" axon_status axon_indeg(axon_net* n, uint32_t* out) { return per_neuron(n, out, axon::launch_indeg); }
extern "C" axon_status axon_untouched(axon_net* n, uint32_t* out) { return per_neuron(n, out, axon::launch_untouched); }

static axon_status wdiag(axon_net* n, const uint32_t* memb, const uint32_t* asm_spk, uint32_t nstim,
                         uint32_t sel, uint32_t hist[48]) {
    if (!n || !memb || !asm_spk || !hist) return fail(AXON_ERR_INVALID_ARG, "null argument");
    if (nstim > 32) return fail(AXON_ERR_INVALID_ARG, "nstim must be <= 32 (membership is a u32 mask)");
    HIPCHECK(hipSetDevice(n->device));
    uint32_t* d_asm = nullptr;
    const size_t asmBytes = std::max<size_t>((size_t)nstim * n->NCOL * 4, 4);
    HIPCHECK(hipMalloc(&d_asm, asmBytes));
    hipError_t e = hipMemcpyAsync(n->d_mask, memb, (size_t)n->N * 4, hipMemcpyHostToDevice, n->stream);
    if (e == hipSuccess && nstim) e = hipMemcpyAsync(d_asm, asm_spk, (size_t)nstim * n->NCOL * 4, hipMemcpyHostToDevice, n->stream);
    if (e == hipSuccess) e = hipMemsetAsync(n->d_small, 0, 48 * 4, n->stream);
    if (e == hipSuccess) e = axon::launch_wdiag(n->d, n->d_mask, d_asm, nstim, sel, n->d_ball, n->d_small, n->stream);
    uint32_t tmp[48];
    if (e == hipSuccess) e = hipMemcpyAsync(tmp, n->d_small, 48 * 4, hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    (void)hipFree(d_asm);
    if (e != hipSuccess) return fail(AXON_ERR_DEVICE, "wdiag: %s", hipGetErrorString(e));
    memcpy(hist, tmp, sizeof tmp);  // zeroed before each call (mm:1127)
    return AXON_OK;
}

// ================================================================== kernel harnesses
namespace {
struct DevTmp {
    std::vector<void*> ps;
    ~DevTmp() { for (void* p : ps) (void)hipFree(p); }
    hipError_t alloc(uint32_t** p, size_t bytes) {
        hipError_t e = hipMalloc((void**)p, std::max<size_t>(bytes, 4));
        if (e == hipSuccess) ps.push_back(*p);
        return e;
    }
};
}  // namespace

extern "C" axon_status axon_bs_test(uint32_t op, uint32_t lo, uint32_t hi, uint32_t* a0, uint32_t* a1,
                                    uint32_t* a2, uint32_t* a3, const uint32_t* m, size_t nw) {
    if (!a0 || !a1 || !a2 || !a3 || !m) return fail(AXON_ERR_INVALID_ARG, "null argument");
    int dev;
    axon_status st = select_device(&dev);
    if (st != AXON_OK) return st;
    DevTmp t;
    uint32_t* d[5];
    for (int i = 0; i < 5; ++i) HIPCHECK(t.alloc(&d[i], nw * 4));
    uint32_t* h[4] = {a0, a1, a2, a3};
    for (int i = 0; i < 4; ++i) HIPCHECK(hipMemcpy(d[i], h[i], nw * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d[4], m, nw * 4, hipMemcpyHostToDevice));
    HIPCHECK(axon::launch_bstest(op, lo, hi, d[0], d[1], d[2], d[3], d[4], nw, nullptr));
    HIPCHECK(hipDeviceSynchronize());
    for (int i = 0; i < 4; ++i) HIPCHECK(hipMemcpy(h[i], d[i], nw * 4, hipMemcpyDeviceToHost));
    return AXON_OK;
}

extern "C" axon_status axon_kwta_test(const uint32_t* v, uint32_t ncols, uint32_t k, uint32_t* ball, uint32_t* vk,
                                      uint32_t* vk1) {
    if (!v || !ball || !vk || !vk1) return fail(AXON_ERR_INVALID_ARG, "null argument");
    int dev;
    axon_status st = select_device(&dev);
    if (st != AXON_OK) return st;
    DevTmp t;
    uint32_t *dv, *db, *dk, *dk1;
    HIPCHECK(t.alloc(&dv, (size_t)ncols * 32 * 4));
    HIPCHECK(t.alloc(&db, (size_t)ncols * 4));
    HIPCHECK(t.alloc(&dk, (size_t)ncols * 4));
    HIPCHECK(t.alloc(&dk1, (size_t)ncols * 4));
    HIPCHECK(hipMemcpy(dv, v, (size_t)ncols * 32 * 4, hipMemcpyHostToDevice));
    HIPCHECK(axon::launch_kwta_test(dv, ncols, k, db, dk, dk1, nullptr));
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(ball, db, (size_t)ncols * 4, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(vk, dk, (size_t)ncols * 4, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(vk1, dk1, (size_t)ncols * 4, hipMemcpyDeviceToHost));
    return AXON_OK;
}

extern "C" axon_status axon_topk_test(const uint32_t* drive, uint32_t ninst, uint32_t ncol, uint32_t incols,
                                      uint32_t target, uint32_t par, uint32_t* theta) {
    if (!drive || !theta || ninst == 0) return fail(AXON_ERR_INVALID_ARG, "null argument");
    int dev;
    axon_status st = select_device(&dev);
    if (st != AXON_OK) return st;
    DevTmp t;
    uint32_t *dd, *dt;
    HIPCHECK(t.alloc(&dd, (size_t)ninst * ncol * 4));
    HIPCHECK(t.alloc(&dt, (size_t)ninst * 2 * 4));
    HIPCHECK(hipMemcpy(dd, drive, (size_t)ninst * ncol * 4, hipMemcpyHostToDevice));
    HIPCHECK(axon::launch_topk_test(dd, ninst, ncol, incols, target, par, dt, nullptr));
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(theta, dt, (size_t)ninst * 2 * 4, hipMemcpyDeviceToHost));
    return AXON_OK;
}

// ================================================================== probe batches
// probe tiles (probes sharing one propagation wave), in the autotuner's order
static const uint32_t kProbeTiles[AXON_PROBE_TILES] = {1, 2, 4, 8, 16, 32, 64};

// waves per CU the probe propagation may keep resident, in the autotuner's order (0 = as many as fit)
static const uint32_t kProbeWaves[AXON_PROBE_OCCUPANCIES] = {0, 32, 16};

// The readout parameters of a probe batch on network n (present() with
// learn = 0, mm:818-822), and its accumulator width: a neuron's accumulator
// is at most 32 branches x (256 presynaptic bits + boost) in mode 1 (x 256 in
// mode 0, x 1 in mode 2) plus the facilitation, before the bias is taken
// off; when that bound is below 2^16 the probe propagation stores 16-bit
// accumulators, which hold every value exactly.
static void probe_params(const axon_net* n, axon::ProbeDev& d) {
    d.k = n->kTest ? n->kTest : n->K;  // readout k (mm:818)
    d.facil = n->facil;
    d.dthr = n->P.dthr; d.boost = n->P.boost; d.mode = n->P.mode; d.biasShift = n->P.biasShift;
    d.target = n->P.targetCols; d.topkpar = n->d.topkpar;
    const uint64_t perBranch = d.mode == 0u ? 256u : d.mode == 1u ? 256ull + d.boost : 1u;
    const uint64_t bound = (uint64_t)axon::NSLOT * perBranch + d.facil;
    d.acc16 = bound <= 0xFFFFu ? 1u : 0u;
}

struct axon_probes {
    axon_net* net = nullptr;
    axon::ProbeDev d{};
    uint32_t tile = 32;
    uint32_t waves = 0;  // waves per CU of the propagation, 0 = no limit
    std::vector<void*> allocs;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, hipGraphExec_t> graphs;  // (steps, slots advanced, waves)
};

extern "C" void axon_probes_destroy(axon_probes* b) {
    if (!b) return;
    if (b->net && b->net->stream) (void)hipStreamSynchronize(b->net->stream);
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    for (void* p : b->allocs) (void)hipFree(p);
    delete b;
}

static axon_status probes_create(axon_net* n, uint32_t slots, uint32_t max_t, uint32_t tile, axon_probes** out) {
    if (!n || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    bool known = false;
    for (uint32_t t : kProbeTiles) known |= t == tile;
    if (!known) return fail(AXON_ERR_INVALID_ARG, "probe tile must be 1, 2, 4, 8, 16, 32 or 64 (got %u)", tile);
    if (slots == 0 || max_t == 0 || max_t > 4096) return fail(AXON_ERR_INVALID_ARG, "slots >= 1, max_t in 1..4096");
    HIPCHECK(hipSetDevice(n->device));
    axon_probes* b = new axon_probes();
    b->net = n;
    b->tile = tile;
    const uint32_t ns = (slots + tile - 1) / tile * tile;
    const size_t NC = n->NCOL, NN = n->N;
    axon::ProbeDev& d = b->d;
    d.W3 = n->d.W3; d.conn = n->d.conn; d.bias = n->d.bias;
    d.nslots = ns; d.ncol = n->NCOL; d.incols = n->cfg.incols; d.maxt = max_t;
    d.gw = (n->NCOL + 31) / 32;
    probe_params(n, d);
    struct { uint32_t** p; size_t words; } bufs[] = {
        {&d.spk, ns * NC}, {&d.gact, (size_t)ns * d.gw}, {&d.acc, ns * NN}, {&d.drive, ns * NC},
        {&d.stim, ns * NC}, {&d.hist, (size_t)ns * max_t * NC}, {&d.theta, (size_t)ns * 2},
        {&d.st_t, ns}, {&d.st_T, ns}, {&d.st_clamp, ns}, {&d.st_bias, ns}, {&d.selDone, ns}};
    for (auto& x : bufs) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, x.words * 4);
        if (e != hipSuccess) {
            axon_probes_destroy(b);
            return fail(AXON_ERR_OUT_OF_MEMORY, "probe batch: hipMalloc(%zu): %s", x.words * 4, hipGetErrorString(e));
        }
        b->allocs.push_back(p);
        *x.p = (uint32_t*)p;
        if (hipMemsetAsync(p, 0, x.words * 4, n->stream) != hipSuccess) {
            axon_probes_destroy(b);
            return fail(AXON_ERR_DEVICE, "hipMemsetAsync failed");
        }
    }
    if (hipStreamSynchronize(n->stream) != hipSuccess) {
        axon_probes_destroy(b);
        return fail(AXON_ERR_DEVICE, "probe batch init failed");
    }
    *out = b;
    return AXON_OK;
}

extern "C" axon_status axon_probes_set_occupancy(axon_probes* b, uint32_t waves_per_cu) {
    if (!b) return fail(AXON_ERR_INVALID_ARG, "null batch");
    bool known = false;
    for (uint32_t w : kProbeWaves) known |= w == waves_per_cu;
    if (!known) return fail(AXON_ERR_INVALID_ARG, "waves per CU must be 0 (no limit), 32 or 16 (got %u)", waves_per_cu);
    b->waves = waves_per_cu;
    return AXON_OK;
}

extern "C" axon_status axon_probes_info(const axon_probes* b, uint32_t* slots, uint32_t* max_t, uint32_t* tile) {
    if (!b) return fail(AXON_ERR_INVALID_ARG, "null batch");
    if (slots) *slots = b->d.nslots;
    if (max_t) *max_t = b->d.maxt;
    if (tile) *tile = b->tile;
    return AXON_OK;
}

extern "C" axon_status axon_probes_load(axon_probes* b, uint32_t first, uint32_t count, const uint32_t* stims,
                                        const uint32_t* T, uint32_t clamp_steps, uint32_t use_bias) {
    if (!b || !stims || !T) return fail(AXON_ERR_INVALID_ARG, "null argument");
    if (first + count > b->d.nslots || first + count < first)
        return fail(AXON_ERR_INVALID_ARG, "slots [%u,%u) beyond %u", first, first + count, b->d.nslots);
    if (use_bias > 1) return fail(AXON_ERR_INVALID_ARG, "use_bias must be 0 or 1 for probes");
    for (uint32_t i = 0; i < count; ++i)
        if (T[i] > b->d.maxt) return fail(AXON_ERR_CAPACITY, "T=%u exceeds the batch's max_t=%u", T[i], b->d.maxt);
    if (!count) return AXON_OK;
    axon_net* n = b->net;
    HIPCHECK(hipSetDevice(n->device));
    const size_t NC = b->d.ncol;
    std::vector<uint32_t> cl(count, clamp_steps), ub(count, use_bias);
    HIPCHECK(hipMemcpyAsync(b->d.stim + first * NC, stims, count * NC * 4, hipMemcpyHostToDevice, n->stream));
    HIPCHECK(hipMemcpyAsync(b->d.st_T + first, T, count * 4, hipMemcpyHostToDevice, n->stream));
    HIPCHECK(hipMemcpyAsync(b->d.st_clamp + first, cl.data(), count * 4, hipMemcpyHostToDevice, n->stream));
    HIPCHECK(hipMemcpyAsync(b->d.st_bias + first, ub.data(), count * 4, hipMemcpyHostToDevice, n->stream));
    HIPCHECK(axon::launch_probe_reset(b->d, first, count, n->stream));
    // the host vectors above must outlive the copies
    HIPCHECK(hipStreamSynchronize(n->stream));
    return AXON_OK;
}

// Advances slots [0, nact) by `steps` timesteps; every slot from nact on must
// be idle (finished or never admitted), so launches cover only the probe
// tiles that can hold live work.  nact is rounded up to whole tiles.
static axon_status probes_advance(axon_probes* b, uint32_t steps, uint32_t nact) {
    if (!steps || !nact) return AXON_OK;
    axon_net* n = b->net;
    nact = std::min((nact + b->tile - 1u) / b->tile * b->tile, b->d.nslots);
    HIPCHECK(hipSetDevice(n->device));
    if (!n->use_graphs) {
        for (uint32_t i = 0; i < steps; ++i) HIPCHECK(axon::launch_probe_step(b->d, b->tile, nact, b->waves, n->stream));
        return AXON_OK;
    }
    const auto key = std::make_tuple(steps, nact, b->waves);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        hipGraph_t g = nullptr;
        HIPCHECK(hipStreamBeginCapture(n->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (uint32_t i = 0; i < steps && e == hipSuccess; ++i)
            e = axon::launch_probe_step(b->d, b->tile, nact, b->waves, n->stream);
        hipError_t e2 = hipStreamEndCapture(n->stream, &g);
        if (e != hipSuccess) return fail(AXON_ERR_DEVICE, "probe capture: %s", hipGetErrorString(e));
        if (e2 != hipSuccess) return fail(AXON_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e2));
        hipGraphExec_t ex = nullptr;
        hipError_t e3 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e3 != hipSuccess) return fail(AXON_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e3));
        it = b->graphs.emplace(key, ex).first;
    }
    HIPCHECK(hipGraphLaunch(it->second, n->stream));
    return AXON_OK;
}

extern "C" axon_status axon_probes_step(axon_probes* b, uint32_t steps) {
    if (!b) return fail(AXON_ERR_INVALID_ARG, "null batch");
    return probes_advance(b, steps, b->d.nslots);
}

extern "C" axon_status axon_probes_sync(axon_probes* b) {
    if (!b) return fail(AXON_ERR_INVALID_ARG, "null batch");
    HIPCHECK(hipStreamSynchronize(b->net->stream));
    return AXON_OK;
}

extern "C" axon_status axon_probes_read(axon_probes* b, uint32_t first, uint32_t count, uint32_t* spikes,
                                        uint32_t* hist, uint32_t* steps_done) {
    if (!b) return fail(AXON_ERR_INVALID_ARG, "null batch");
    if (first + count > b->d.nslots || first + count < first) return fail(AXON_ERR_INVALID_ARG, "slot range");
    axon_net* n = b->net;
    HIPCHECK(hipSetDevice(n->device));
    HIPCHECK(hipStreamSynchronize(n->stream));
    const size_t NC = b->d.ncol;
    if (spikes) HIPCHECK(hipMemcpy(spikes, b->d.spk + first * NC, count * NC * 4, hipMemcpyDeviceToHost));
    if (hist)
        HIPCHECK(hipMemcpy(hist, b->d.hist + (size_t)first * b->d.maxt * NC, (size_t)count * b->d.maxt * NC * 4,
                           hipMemcpyDeviceToHost));
    if (steps_done) HIPCHECK(hipMemcpy(steps_done, b->d.st_t + first, count * 4, hipMemcpyDeviceToHost));
    return AXON_OK;
}

static axon_status probes_run(axon_probes* b, uint32_t count, const uint32_t* stims, uint32_t T,
                              uint32_t clamp_steps, uint32_t use_bias, uint32_t* hist, uint32_t* spikes, double* ms) {
    if (!b || !stims) return fail(AXON_ERR_INVALID_ARG, "null argument");
    if (count > b->d.nslots) return fail(AXON_ERR_CAPACITY, "%u probes exceed %u slots", count, b->d.nslots);
    std::vector<uint32_t> Ts(b->d.nslots, 0);
    std::fill(Ts.begin(), Ts.begin() + count, T);
    std::vector<uint32_t> zero((size_t)(b->d.nslots - count) * b->d.ncol, 0);
    axon_status st = axon_probes_load(b, 0, count, stims, Ts.data(), clamp_steps, use_bias);
    if (st == AXON_OK && count < b->d.nslots)  // idle slots: T = 0
        st = axon_probes_load(b, count, b->d.nslots - count, zero.data(), Ts.data() + count, clamp_steps, use_bias);
    if (st != AXON_OK) return st;
    const auto t0 = std::chrono::steady_clock::now();
    if ((st = axon_probes_step(b, T)) != AXON_OK) return st;
    if ((st = axon_probes_sync(b)) != AXON_OK) return st;
    if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (hist) {  // rows 0..T-1 of each probe, packed
        axon_net* n = b->net;
        const size_t NC = b->d.ncol;
        for (uint32_t i = 0; i < count; ++i)
            HIPCHECK(hipMemcpy(hist + (size_t)i * T * NC, b->d.hist + (size_t)i * b->d.maxt * NC, (size_t)T * NC * 4,
                               hipMemcpyDeviceToHost));
        (void)n;
    }
    if (spikes) return axon_probes_read(b, 0, count, spikes, nullptr, nullptr);
    return AXON_OK;
}

// The probe run of axon_probes_run with every piece of slot state carved out
// of the library's static per-GPU scratch array instead of a probe batch's
// allocations: it allocates nothing, captures no graph and issues no memset,
// so a caller may run it under stream capture.  All calls on one GPU share
// the array; the setup kernel clears what must start at zero.
extern "C" axon_status axon_probes_run_noalloc(axon_net* n, uint32_t count, const uint32_t* d_stims, uint32_t T,
                                               uint32_t clamp_steps, uint32_t use_bias, uint32_t tile,
                                               uint32_t* d_hist, uint32_t* d_spikes, void* stream) {
    if (!n || !d_stims) return fail(AXON_ERR_INVALID_ARG, "null argument");
    bool known = false;
    for (uint32_t t : kProbeTiles) known |= t == tile;
    if (!known) return fail(AXON_ERR_INVALID_ARG, "probe tile must be 1, 2, 4, 8, 16, 32 or 64 (got %u)", tile);
    if (count == 0 || T == 0 || T > 4096) return fail(AXON_ERR_INVALID_ARG, "count >= 1 and T in 1..4096");
    if (use_bias > 1) return fail(AXON_ERR_INVALID_ARG, "use_bias must be 0 or 1 for probes");
    const size_t NC = n->NCOL, NN = n->N, gw = (n->NCOL + 31) / 32;
    const size_t ns = ((size_t)count + tile - 1) / tile * tile;
    auto al = [](size_t w) { return (w + 63) & ~(size_t)63; };  // 256-byte aligned pieces
    const size_t per[] = {ns * NC, ns * gw, ns * NN, ns * NC, 2 * ns, ns, ns, ns, ns, ns,
                          d_hist ? 0 : ns * (size_t)T * NC};
    size_t need = 0;
    for (size_t w : per) need += al(w);
    if (need > axon::NOALLOC_SCRATCH_WORDS)
        return fail(AXON_ERR_CAPACITY,
                    "allocation-free probe run: %u probes (%zu slots at tile %u) of ncol=%u need %zu bytes of "
                    "state, the static buffer holds %zu",
                    count, ns, tile, n->NCOL, need * 4, axon::NOALLOC_SCRATCH_WORDS * 4);
    HIPCHECK(hipSetDevice(n->device));
    uint32_t* base = nullptr;
    HIPCHECK(axon::noalloc_scratch(n->device, &base));
    uint32_t* cur = base;
    auto take = [&](size_t w) { uint32_t* p = cur; cur += al(w); return p; };
    axon::ProbeDev d{};
    d.W3 = n->d.W3; d.conn = n->d.conn; d.bias = n->d.bias;
    d.spk = take(per[0]); d.gact = take(per[1]); d.acc = take(per[2]); d.drive = take(per[3]);
    d.theta = take(per[4]); d.st_t = take(per[5]); d.st_T = take(per[6]); d.st_clamp = take(per[7]);
    d.st_bias = take(per[8]); d.selDone = take(per[9]);
    d.hist = d_hist ? d_hist : take(per[10]);
    d.stim = const_cast<uint32_t*>(d_stims);  // read only: slots below count, during their clamped steps
    d.nslots = (uint32_t)ns; d.ncol = n->NCOL; d.incols = n->cfg.incols; d.maxt = T; d.gw = (uint32_t)gw;
    probe_params(n, d);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : n->stream;
    HIPCHECK(axon::launch_probe_setup(d, count, T, clamp_steps, use_bias, s));
    for (uint32_t t = 0; t < T; ++t) HIPCHECK(axon::launch_probe_step(d, tile, d.nslots, 0u, s));
    if (d_spikes) HIPCHECK(axon::launch_copy_words(d_spikes, d.spk, (size_t)count * NC, s));
    return AXON_OK;
}

// Recall cue with replacement (an extension of loadStim, mm:800-804): start
// from pattern p's stimulus columns, replace `nreplace` of them by input
// columns outside the pattern.  Both choices are partial Fisher-Yates draws
// with the libc++ uniform index over mt19937(seed).  pattern = 0xFFFFFFFF
// draws `stimcols` input columns uniformly instead.
extern "C" axon_status axon_probe_cue(const axon_net* n, uint32_t pattern, uint32_t nreplace, uint32_t seed,
                                      uint32_t* stim) {
    if (!n || !stim) return fail(AXON_ERR_INVALID_ARG, "null argument");
    const uint32_t IN = n->cfg.incols, SC = n->cfg.stimcols;
    std::mt19937 rng(seed);
    memset(stim, 0, (size_t)n->NCOL * 4);
    auto pick = [&](std::vector<uint32_t>& v, uint32_t r) {  // first r entries become a uniform r-subset
        for (uint32_t i = 0; i < r && i + 1 < v.size(); ++i) {
            const uint64_t j = axon::libcxx_uniform_index(rng, (uint64_t)(v.size() - i));
            std::swap(v[i], v[i + j]);
        }
    };
    if (pattern == 0xFFFFFFFFu) {
        std::vector<uint32_t> all(IN);
        for (uint32_t i = 0; i < IN; ++i) all[i] = i;
        pick(all, SC);
        for (uint32_t i = 0; i < SC; ++i) stim[all[i]] = 0xFFFFFFFFu;
        return AXON_OK;
    }
    if (pattern >= n->stimCols.size()) return fail(AXON_ERR_INVALID_ARG, "pattern %u out of range", pattern);
    std::vector<uint32_t> mem = n->stimCols[pattern];
    if (nreplace > mem.size() || nreplace > IN - mem.size())
        return fail(AXON_ERR_INVALID_ARG, "cannot replace %u of %zu columns", nreplace, mem.size());
    std::vector<uint8_t> isMem(IN, 0);
    for (uint32_t c : mem) isMem[c] = 1;
    std::vector<uint32_t> non;
    for (uint32_t c = 0; c < IN; ++c)
        if (!isMem[c]) non.push_back(c);
    pick(mem, nreplace);
    pick(non, nreplace);
    for (uint32_t i = nreplace; i < mem.size(); ++i) stim[mem[i]] = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < nreplace; ++i) stim[non[i]] = 0xFFFFFFFFu;
    return AXON_OK;
}

// ================================================================== many networks
// Networks are trained in up to MULTI_GROUPS lock-step groups, one stream
// each, and no more groups than the process has hardware queues
// (GPU_MAX_HW_QUEUES, the HIP runtime's setting, 4 by default): networks with
// the same queue length share a group (merged down to the limit, shortest
// first), so a group's launches never carry idle networks for long, and the
// groups' kernels overlap on the device (one group's latency-bound top-K runs
// beside another's propagation).  Measured on 25 networks (4 hardware
// queues): 3 groups 0.925 s, 4 groups 0.945 s; with as many queues as
// groups, 2: 0.97 s, 3: 0.93 s, 4: 0.94 s, 5: 1.16 s, 8: 2.1 s.
constexpr uint32_t MULTI_GROUPS = 3;
static uint32_t multi_groups() {
    const char* e = getenv("GPU_MAX_HW_QUEUES");
    const int q = e ? atoi(e) : (int)MULTI_GROUPS;
    return (uint32_t)std::min<int>(std::max(q, 1), (int)MULTI_GROUPS);
}

struct MultiGroup {
    axon::MultiDev m{};
    hipGraphExec_t graph = nullptr;
    size_t rounds = 0;
    uint32_t stream = 0;
};

struct axon_multi {
    std::vector<axon_net*> nets;
    std::vector<NetDev> devs;      // host copies of the networks' descriptors
    std::vector<uint64_t> tabs;    // device clamp-pattern tables, one per network
    hipStream_t streams[MULTI_GROUPS] = {};
    std::vector<void*> allocs;     // clamp tables (lifetime of the batch)
    uint32_t ncol = 0, nstats = 0;
    Params P{};
    uint32_t T = 0, trainClamp = 0, contrast = 0, ltdEvery = 1;
};

extern "C" void axon_multi_destroy(axon_multi* b) {
    if (!b) return;
    for (hipStream_t s : b->streams)
        if (s) (void)hipStreamSynchronize(s);
    for (void* p : b->allocs) (void)hipFree(p);
    for (hipStream_t s : b->streams)
        if (s) (void)hipStreamDestroy(s);
    delete b;
}

// Networks must share every dynamics parameter; seeds, stored-pattern counts
// and epochs may differ.
static axon_status multi_create(axon_net* const* nets, uint32_t nnet, axon_multi** out) {
    if (!nets || !out || nnet == 0) return fail(AXON_ERR_INVALID_ARG, "null argument or no networks");
    *out = nullptr;
    const axon_config& c0 = nets[0]->cfg;
    for (uint32_t i = 0; i < nnet; ++i) {
        if (!nets[i]) return fail(AXON_ERR_INVALID_ARG, "network %u is null", i);
        const axon_config& c = nets[i]->cfg;
        axon_config a = c, b = c0;
        a.seed = b.seed = 0; a.stim = b.stim = 0; a.epochs = b.epochs = 0;
        if (memcmp(&a, &b, sizeof a) != 0 || nets[i]->d.topkpar != nets[0]->d.topkpar)
            return fail(AXON_ERR_INVALID_ARG, "network %u differs from network 0 in a dynamics parameter", i);
        for (uint32_t j = 0; j < i; ++j)
            if (nets[j] == nets[i]) return fail(AXON_ERR_INVALID_ARG, "network %u appears twice", i);
    }
    HIPCHECK(hipSetDevice(nets[0]->device));
    axon_multi* b = new axon_multi();
    b->nets.assign(nets, nets + nnet);
    b->ncol = nets[0]->NCOL;
    b->nstats = nets[0]->NSTATS;
    b->P = nets[0]->P;
    b->T = c0.T;
    b->trainClamp = c0.trainclamp;
    b->contrast = c0.contrast;
    b->ltdEvery = nets[0]->ltdEvery;
    auto bail = [&](axon_status s) { axon_multi_destroy(b); return s; };
    for (hipStream_t& st : b->streams)
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
            return bail(fail(AXON_ERR_DEVICE, "hipStreamCreate failed"));
    b->devs.resize(nnet);
    b->tabs.resize(nnet);
    for (uint32_t i = 0; i < nnet; ++i) {
        axon_net* n = nets[i];
        if (hipStreamSynchronize(n->stream) != hipSuccess) return bail(fail(AXON_ERR_DEVICE, "net stream"));
        b->devs[i] = n->d;
        // clamp patterns at keep = 1.0: every stimulus column of the pattern (loadStim, mm:800-804)
        std::vector<uint32_t> tab((size_t)n->stimCols.size() * b->ncol, 0u);
        for (size_t q = 0; q < n->stimCols.size(); ++q)
            for (uint32_t col : n->stimCols[q]) tab[q * b->ncol + col] = 0xFFFFFFFFu;
        void* p = nullptr;
        if (hipMalloc(&p, tab.size() * 4) != hipSuccess) return bail(fail(AXON_ERR_OUT_OF_MEMORY, "clamp table"));
        b->allocs.push_back(p);
        if (hipMemcpy(p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(AXON_ERR_DEVICE, "stim table copy"));
        b->tabs[i] = (uint64_t)(uintptr_t)p;
    }
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(AXON_ERR_DEVICE, "stim table copy"));
    *out = b;
    return AXON_OK;
}

// One presentation of every scheduled network of a group (mm:837-884 with
// learn = 1), followed by the group's round counter advance.  As in
// enqueue_steps, step t's plasticity runs inside step t + 1's propagation and
// the last step's in a launch of its own.
static hipError_t multi_enqueue_round(const axon_multi* b, const axon::MultiDev& m, hipStream_t s) {
    Params P = b->P;
    P.facil = 0u;
    P.useBias = 1u;
    const uint32_t T = b->T, clampSteps = b->trainClamp;
    // with two or more steps, step 0 (spikes start cleared) is one launch that also does the begin work,
    // or, when clamped, runs inside step 1's propagation
    const bool first = T >= 2u, step0 = first && clampSteps >= 1u;
    hipError_t e = first ? hipSuccess : axon::launch_multi_begin(m, b->ncol, b->nstats, s);
    Params PL{}, S0{};
    bool pending = false;
    for (uint32_t t = 0; t < T && e == hipSuccess; ++t) {
        P.step = t;
        P.learn = 1u;
        P.freeze = 0u;
        P.clamp = (t < clampSteps) ? 1u : 0u;
        P.depress = (t % 2u == 1u) ? 1u : 0u;
        P.phase = !b->contrast ? 0u : (t < clampSteps) ? 1u : (((t - clampSteps) % b->ltdEvery == 0u) ? 2u : 3u);
        if (first && t == 0u) {
            if (step0) S0 = P;
            else e = axon::launch_multi_emit_first(m, P, true, b->ncol, s);
            continue;
        }
        e = axon::launch_multi_drive(m, P, t, b->ncol, s, pending ? &PL : nullptr, step0 && t == 1u ? &S0 : nullptr);
        if (e == hipSuccess) e = axon::launch_multi_topk(m, P, s);
        pending = P.phase != 3u;
        PL = P;
        if (e == hipSuccess && pending && t + 1u == T) {
            // the last step's plasticity with its emit, a launch that also ends the round
            e = axon::launch_multi_emit_learn(m, P, t, b->ncol, s);
            return e;
        }
        if (e == hipSuccess) e = axon::launch_multi_emit(m, P, t, true, b->ncol, s);
    }
    // without plasticity at the last step, k_mb_end ends the round
    if (e == hipSuccess) e = axon::launch_multi_end(m, s);
    return e;
}

// Trains every network with the reference protocol: epochs[i] x stim_i
// presentations (pattern k mod stim_i, keep 1.0; mm:1009-1010), then, when
// extra = 1, the additional presentation of pattern 0 (mm:1033).  Network i's
// work queue holds its own presentations; within a group all networks
// advance one presentation per round and drop out when their queue is empty.
static axon_status multi_train(axon_multi* b, const uint32_t* epochs, uint32_t extra, double* ms) {
    if (!b || !epochs) return fail(AXON_ERR_INVALID_ARG, "null argument");
    const uint32_t nnet = (uint32_t)b->nets.size();
    std::vector<uint32_t> len(nnet);
    for (uint32_t i = 0; i < nnet; ++i)
        len[i] = epochs[i] * (uint32_t)b->nets[i]->stimCols.size() + (extra ? 1u : 0u);
    // groups: equal queue lengths together, longest first; merge the shortest down to the limit
    std::vector<std::vector<uint32_t>> members;
    {
        std::vector<uint32_t> order(nnet);
        for (uint32_t i = 0; i < nnet; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return len[x] > len[y]; });
        for (uint32_t i : order) {
            if (len[i] == 0u) continue;
            if (members.empty() || len[members.back()[0]] != len[i]) members.emplace_back();
            members.back().push_back(i);
        }
        const uint32_t ng = multi_groups();
        while (members.size() > ng) {
            auto last = members.back();
            members.pop_back();
            members.back().insert(members.back().end(), last.begin(), last.end());
        }
    }
    HIPCHECK(hipSetDevice(b->nets[0]->device));
    for (axon_net* n : b->nets) HIPCHECK(hipStreamSynchronize(n->stream));

    std::vector<MultiGroup> groups(members.size());
    std::vector<void*> tmp;  // this call's descriptors, schedules and counters
    auto release = [&]() {
        for (hipStream_t s : b->streams) (void)hipStreamSynchronize(s);
        for (MultiGroup& g : groups)
            if (g.graph) (void)hipGraphExecDestroy(g.graph);
        for (void* p : tmp) (void)hipFree(p);
    };
    auto dalloc = [&](size_t bytes, const void* src) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        tmp.push_back(p);
        if (src && hipMemcpy(p, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        if (!src && hipMemset(p, 0, bytes) != hipSuccess) return nullptr;
        return p;
    };
    const bool graphs = b->nets[0]->use_graphs;
    size_t maxRounds = 0;
    for (size_t gi = 0; gi < members.size(); ++gi) {
        MultiGroup& g = groups[gi];
        const std::vector<uint32_t>& mem = members[gi];
        const uint32_t k = (uint32_t)mem.size();
        for (uint32_t i : mem) g.rounds = std::max<size_t>(g.rounds, len[i]);
        maxRounds = std::max(maxRounds, g.rounds);
        std::vector<NetDev> devs(k);
        std::vector<uint64_t> tabs(k);
        std::vector<uint32_t> sched(g.rounds * k, 0xFFFFFFFFu);
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t i = mem[j];
            devs[j] = b->devs[i];
            tabs[j] = b->tabs[i];
            const uint32_t ns = (uint32_t)b->nets[i]->stimCols.size(), body = len[i] - (extra ? 1u : 0u);
            for (uint32_t r = 0; r < len[i]; ++r) sched[(size_t)r * k + j] = r < body ? r % ns : 0u;
        }
        g.m.nets = (const NetDev*)dalloc(sizeof(NetDev) * k, devs.data());
        g.m.stimTabs = (const uint64_t*)dalloc(8 * k, tabs.data());
        g.m.sched = (const uint32_t*)dalloc(sched.size() * 4, sched.data());
        g.m.round = (uint32_t*)dalloc(4, nullptr);
        g.m.nnet = k;
        g.stream = (uint32_t)gi;
        if (!g.m.nets || !g.m.stimTabs || !g.m.sched || !g.m.round) {
            release();
            return fail(AXON_ERR_OUT_OF_MEMORY, "multi-network descriptors");
        }
        if (!graphs) continue;
        hipStream_t s = b->streams[gi];
        hipGraph_t gr = nullptr;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            release();
            return fail(AXON_ERR_DEVICE, "hipStreamBeginCapture failed");
        }
        hipError_t e = multi_enqueue_round(b, g.m, s);
        hipError_t e2 = hipStreamEndCapture(s, &gr);
        hipError_t e3 = (e == hipSuccess && e2 == hipSuccess) ? hipGraphInstantiate(&g.graph, gr, nullptr, nullptr, 0)
                                                             : hipErrorUnknown;
        if (gr) (void)hipGraphDestroy(gr);
        if (e3 != hipSuccess) {
            release();
            return fail(AXON_ERR_DEVICE, "multi capture: %s / %s", hipGetErrorString(e), hipGetErrorString(e2));
        }
    }
    // the descriptor uploads and the round counters' fills above went through
    // the null stream, which the groups' non-blocking streams do not wait for
    if (hipDeviceSynchronize() != hipSuccess) {
        release();
        return fail(AXON_ERR_DEVICE, "multi train: upload");
    }
    const auto t0 = std::chrono::steady_clock::now();
    hipError_t e = hipSuccess;
    for (size_t r = 0; r < maxRounds && e == hipSuccess; ++r)
        for (MultiGroup& g : groups) {
            if (r >= g.rounds) continue;
            hipStream_t s = b->streams[g.stream];
            e = graphs ? hipGraphLaunch(g.graph, s) : multi_enqueue_round(b, g.m, s);
            if (e != hipSuccess) break;
        }
    for (hipStream_t s : b->streams)
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    const auto t1 = std::chrono::steady_clock::now();
    release();
    if (e != hipSuccess) return fail(AXON_ERR_DEVICE, "multi train: %s", hipGetErrorString(e));
    if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return AXON_OK;
}

// Order-independent 64-bit checksum of an engine buffer (sum over words of
// w * (2i+1) plus a per-word hash), for comparing states without copying them.
extern "C" axon_status axon_checksum(axon_net* n, axon_buffer which, uint64_t* sum) {
    if (!n || !sum) return fail(AXON_ERR_INVALID_ARG, "null argument");
    void* p; size_t words;
    axon_status st = buffer_ptr(n, which, &p, &words);
    if (st != AXON_OK) return st;
    if (which == AXON_BUF_ACC && n->d.acc16) {  // 16-bit accumulators: the sum over their 32-bit values
        std::vector<uint32_t> v(words);
        if ((st = axon_read_buffer(n, which, v.data(), words * 4)) != AXON_OK) return st;
        uint64_t acc = 0;
        for (size_t i = 0; i < words; ++i) {
            uint32_t x = v[i] ^ (uint32_t)i ^ (uint32_t)(i >> 32);  // hash32 (lowbias32), as k_checksum
            x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
            acc += (uint64_t)v[i] * (2ull * i + 1ull) + x;
        }
        *sum = acc;
        return AXON_OK;
    }
    HIPCHECK(hipSetDevice(n->device));
    HIPCHECK(hipMemsetAsync(n->d_small, 0, 8, n->stream));
    HIPCHECK(axon::launch_checksum((const uint32_t*)p, words, (unsigned long long*)n->d_small, n->stream));
    uint64_t v = 0;
    HIPCHECK(hipMemcpyAsync(&v, n->d_small, 8, hipMemcpyDeviceToHost, n->stream));
    HIPCHECK(hipStreamSynchronize(n->stream));
    *sum = v;
    return AXON_OK;
}

// ================================================================== serving
// Continuous batching of recall queries over resident networks.  Each network
// owns a probe batch (its slots); queued requests are admitted in arrival
// order into free slots while the HBM budget allows, every busy slot
// advances together, and finished slots are gathered and freed at once, so
// later requests join while earlier ones are still running.  A slot's
// computation depends only on its own cue and the network, which is what
// makes a result the same alone and in any mix.
//
// Ticks are pipelined two deep.  The host keeps each slot's step count
// itself (every busy slot advances one step per step launched), so it knows
// which queries a tick finishes without reading counters back: tick n
// enqueues, on each network's stream, its admissions, its steps, the gather
// of the slots it finishes and the copy of their results, and only then
// waits for tick n - 1's results.  The device is never left idle while the
// host admits and collects.  A finished slot is free for the next tick's
// admissions at once (its gather precedes them on the stream); its bytes
// leave the budget then too.  Host buffers alternate between ticks.
#include <deque>
#include <queue>
#include <set>
#include <unordered_map>

namespace {
struct SrvReq {
    uint64_t id;
    uint32_t net, T, clamp, bias;
    double arrival;
    uint64_t bytes;
    std::vector<uint32_t> stim;
};
struct SrvSlot {
    bool busy = false;
    uint64_t id = 0, bytes = 0;
    uint32_t T = 0, done = 0;  // steps to run, steps launched
    double arrival = 0;
};
struct SrvResult {
    uint32_t T;
    double arrival, done;
    std::vector<uint32_t> spikes;
    std::vector<uint64_t> digest;
};
struct SrvFinished {
    uint64_t id;
    uint32_t T, off;
    double arrival;
};
// one tick's host and device buffers of a network (two alternate)
struct SrvTick {
    uint32_t *h_list = nullptr, *d_list = nullptr;  // admit entries
    uint32_t *h_g = nullptr, *d_g = nullptr;        // gather slots + offsets
    uint32_t *h_out = nullptr, *d_out = nullptr;    // gathered results
    hipEvent_t ev = nullptr;
    bool pending = false;                           // work enqueued by this tick not yet waited for
    std::vector<SrvFinished> fin;
};
struct SrvNet {
    axon_probes* pb = nullptr;
    std::vector<SrvSlot> slots;
    std::set<uint32_t> free;  // free slots, taken lowest first
    std::deque<SrvReq> queue;
    SrvTick tk[2];
};
}  // namespace

struct axon_server {
    std::vector<axon_net*> nets;
    std::vector<SrvNet> sn;
    axon_server_config cfg{};
    std::unordered_map<uint64_t, SrvResult> done;
    std::vector<uint64_t> last_done;  // ids the last tick collected
    uint64_t next_id = 1, reserved = 0;
    // inflight: admitted and not yet collected; reserved: bytes of the
    // queries still running (a finished query's leave at its gather)
    uint32_t ncol = 0, inflight = 0, queued = 0, cur = 0;
    std::chrono::steady_clock::time_point t0;
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

static bool tune_find(uint32_t ncol, uint32_t probes, uint32_t T, axon_tune_entry* out);

static uint64_t request_bytes(const axon_server* s, uint32_t T) {
    const uint64_t NC = s->ncol;
    return (uint64_t)T * NC * 4 + NC * 32 * 4 + 4 * NC * 4;  // history + accumulators + spikes/drive/stim/gact
}

extern "C" void axon_server_destroy(axon_server* s) {
    if (!s) return;
    for (size_t i = 0; i < s->sn.size(); ++i) {
        SrvNet& v = s->sn[i];
        if (s->nets[i]) (void)hipStreamSynchronize(s->nets[i]->stream);
        for (SrvTick& t : v.tk) {
            for (uint32_t* p : {t.h_list, t.h_g, t.h_out})
                if (p) (void)hipHostFree(p);
            for (uint32_t* p : {t.d_list, t.d_g, t.d_out})
                if (p) (void)hipFree(p);
            if (t.ev) (void)hipEventDestroy(t.ev);
        }
        axon_probes_destroy(v.pb);
    }
    delete s;
}

static axon_status server_create(axon_net* const* nets, uint32_t nnet, const axon_server_config* cfg,
                                 axon_server** out) {
    if (!nets || !nets[0] || !cfg || !out || nnet == 0) return fail(AXON_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (cfg->slots_per_net == 0 || cfg->steps_per_tick == 0) return fail(AXON_ERR_INVALID_ARG, "slots and steps_per_tick >= 1");
    axon_server* s = new axon_server();
    s->cfg = *cfg;
    s->ncol = nets[0]->NCOL;
    if (cfg->tile == 0) {  // the tuning table's choice for this batch, else tile 4 at 32 waves per CU
        axon_tune_entry e{};
        const bool hit = tune_find(s->ncol, cfg->slots_per_net, cfg->max_t, &e);
        s->cfg.tile = hit ? e.tile : 4u;
        s->cfg.waves_per_cu = hit ? e.waves_per_cu : 32u;
    }
    const axon_server_config& C = s->cfg;
    s->nets.assign(nets, nets + nnet);
    s->sn.resize(nnet);
    auto bail = [&](axon_status st) { axon_server_destroy(s); return st; };
    for (uint32_t i = 0; i < nnet; ++i) {
        if (!nets[i] || nets[i]->NCOL != s->ncol) return bail(fail(AXON_ERR_INVALID_ARG, "network %u missing or of another size", i));
        SrvNet& v = s->sn[i];
        axon_status st = axon_probes_create(nets[i], C.slots_per_net, C.max_t, C.tile, &v.pb);
        if (st == AXON_OK) st = axon_probes_set_occupancy(v.pb, C.waves_per_cu);
        if (st != AXON_OK) return bail(st);
        const uint32_t ns = v.pb->d.nslots;
        v.slots.resize(ns);
        for (uint32_t k = 0; k < ns; ++k) v.free.insert(k);
        const size_t listW = (size_t)ns * (4 + s->ncol), outW = (size_t)ns * (s->ncol + 2 * C.max_t);
        for (SrvTick& t : v.tk) {
            if (hipHostMalloc((void**)&t.h_list, listW * 4, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc((void**)&t.h_g, (size_t)ns * 8, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc((void**)&t.h_out, outW * 4, hipHostMallocDefault) != hipSuccess)
                return bail(fail(AXON_ERR_OUT_OF_MEMORY, "pinned host buffers"));
            if (hipMalloc((void**)&t.d_list, listW * 4) != hipSuccess || hipMalloc((void**)&t.d_g, (size_t)ns * 8) != hipSuccess ||
                hipMalloc((void**)&t.d_out, outW * 4) != hipSuccess)
                return bail(fail(AXON_ERR_OUT_OF_MEMORY, "server device buffers"));
            if (hipEventCreateWithFlags(&t.ev, hipEventDisableTiming) != hipSuccess)
                return bail(fail(AXON_ERR_DEVICE, "hipEventCreate failed"));
        }
    }
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return AXON_OK;
}

extern "C" double axon_server_clock(const axon_server* s) { return s ? s->now() : 0.0; }

static axon_status server_submit(axon_server* s, uint32_t net, const uint32_t* stim, uint32_t T,
                                 uint32_t clamp_steps, uint32_t use_bias, double arrival_s, uint64_t* id) {
    if (!s || !stim || !id) return fail(AXON_ERR_INVALID_ARG, "null argument");
    if (net >= s->nets.size()) return fail(AXON_ERR_INVALID_ARG, "network %u out of range", net);
    if (T == 0 || T > s->cfg.max_t) return fail(AXON_ERR_CAPACITY, "T=%u outside 1..max_t=%u", T, s->cfg.max_t);
    if (use_bias > 1) return fail(AXON_ERR_INVALID_ARG, "use_bias must be 0 or 1");
    const uint64_t bytes = request_bytes(s, T);
    if (s->cfg.hbm_budget && bytes > s->cfg.hbm_budget)
        return fail(AXON_ERR_BUDGET_EXCEEDED, "request needs %llu bytes, budget is %llu", (unsigned long long)bytes,
                    (unsigned long long)s->cfg.hbm_budget);
    SrvReq r;
    r.id = s->next_id++;
    r.net = net; r.T = T; r.clamp = clamp_steps; r.bias = use_bias;
    r.arrival = arrival_s >= 0 ? arrival_s : s->now();
    r.bytes = bytes;
    r.stim.assign(stim, stim + s->ncol);
    s->sn[net].queue.push_back(std::move(r));
    s->queued++;
    *id = s->next_id - 1;
    return AXON_OK;
}

// Wait for a tick's work, then record the results it copied back.
static axon_status server_collect(axon_server* s, SrvTick& t, uint32_t& ndone) {
    if (!t.pending) return AXON_OK;
    HIPCHECK(hipEventSynchronize(t.ev));
    t.pending = false;
    const uint32_t NC = s->ncol;
    const double tdone = s->now();
    for (const SrvFinished& f : t.fin) {
        const uint32_t* o = t.h_out + f.off;
        SrvResult r;
        r.T = f.T; r.arrival = f.arrival; r.done = tdone;
        r.spikes.assign(o, o + NC);
        r.digest.resize(f.T);
        for (uint32_t k = 0; k < f.T; ++k) r.digest[k] = (uint64_t)o[NC + 2 * k] | ((uint64_t)o[NC + 2 * k + 1] << 32);
        s->done.emplace(f.id, std::move(r));
        s->last_done.push_back(f.id);
        s->inflight--;
        ++ndone;
    }
    t.fin.clear();
    return AXON_OK;
}

static axon_status server_tick(axon_server* s, uint32_t* completed) {
    if (!s) return fail(AXON_ERR_INVALID_ARG, "null server");
    const uint32_t nnet = (uint32_t)s->nets.size(), NC = s->ncol, spt = s->cfg.steps_per_tick;
    HIPCHECK(hipSetDevice(s->nets[0]->device));
    const uint32_t cur = s->cur, prev = cur ^ 1u;
    // 1. admission in arrival order over the networks' queues: the earliest
    //    request whose network has a free slot goes next; admission stops at
    //    the first such request the HBM budget cannot take
    std::vector<uint32_t> nadm(nnet, 0);
    {
        using Head = std::pair<uint64_t, uint32_t>;  // (request id, network)
        std::priority_queue<Head, std::vector<Head>, std::greater<Head>> heads;
        for (uint32_t i = 0; i < nnet; ++i)
            if (!s->sn[i].queue.empty() && !s->sn[i].free.empty()) heads.emplace(s->sn[i].queue.front().id, i);
        while (!heads.empty()) {
            const uint32_t i = heads.top().second;
            heads.pop();
            SrvNet& v = s->sn[i];
            SrvReq& r = v.queue.front();
            if (s->cfg.hbm_budget && s->reserved + r.bytes > s->cfg.hbm_budget) break;
            const uint32_t slot = *v.free.begin();
            v.free.erase(v.free.begin());
            uint32_t* e = v.tk[cur].h_list + (size_t)nadm[i]++ * (4 + NC);
            e[0] = slot; e[1] = r.T; e[2] = r.clamp; e[3] = r.bias;
            memcpy(e + 4, r.stim.data(), (size_t)NC * 4);
            v.slots[slot] = SrvSlot{true, r.id, r.bytes, r.T, 0u, r.arrival};
            s->reserved += r.bytes;
            s->inflight++;
            s->queued--;
            v.queue.pop_front();
            if (!v.queue.empty() && !v.free.empty()) heads.emplace(v.queue.front().id, i);
        }
    }
    // 2. per network, on its stream: admissions, steps, gather of the slots
    //    these steps finish, the copy of their results, an event
    for (uint32_t i = 0; i < nnet; ++i) {
        SrvNet& v = s->sn[i];
        SrvTick& t = v.tk[cur];
        hipStream_t st = s->nets[i]->stream;
        if (nadm[i]) {
            HIPCHECK(hipMemcpyAsync(t.d_list, t.h_list, (size_t)nadm[i] * (4 + NC) * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(axon::launch_probe_admit(v.pb->d, t.d_list, nadm[i], st));
        }
        uint32_t hi = 0;  // slots are taken lowest-first: launches cover [0, last busy slot]
        for (uint32_t k = 0; k < v.slots.size(); ++k)
            if (v.slots[k].busy) hi = k + 1u;
        if (hi) {
            axon_status r = probes_advance(v.pb, spt, hi);
            if (r != AXON_OK) return r;
            const uint32_t ns = (uint32_t)v.slots.size();
            uint32_t off = 0, nf = 0;
            for (uint32_t k = 0; k < hi; ++k) {
                SrvSlot& sl = v.slots[k];
                if (!sl.busy) continue;
                sl.done = std::min(sl.T, sl.done + spt);
                if (sl.done < sl.T) continue;
                t.h_g[nf] = k;
                t.h_g[ns + nf] = off;
                t.fin.push_back(SrvFinished{sl.id, sl.T, off, sl.arrival});
                off += NC + 2 * sl.T;
                ++nf;
                s->reserved -= sl.bytes;
                sl = SrvSlot{};
                v.free.insert(k);
            }
            if (nf) {
                HIPCHECK(hipMemcpyAsync(t.d_g, t.h_g, (size_t)ns * 8, hipMemcpyHostToDevice, st));
                HIPCHECK(axon::launch_probe_gather(v.pb->d, t.d_g, t.d_g + ns, nf, t.d_out, st));
                HIPCHECK(hipMemcpyAsync(t.h_out, t.d_out, (size_t)off * 4, hipMemcpyDeviceToHost, st));
            }
        }
        // the event also guards this tick's host buffers, which the tick
        // after next fills again (the copies from them are asynchronous)
        if (nadm[i] || hi) {
            HIPCHECK(hipEventRecord(t.ev, st));
            t.pending = true;
        }
    }
    // 3. collect the previous tick's results while this tick runs
    uint32_t ndone = 0;
    s->last_done.clear();
    for (uint32_t i = 0; i < nnet; ++i) {
        axon_status r = server_collect(s, s->sn[i].tk[prev], ndone);
        if (r != AXON_OK) return r;
    }
    s->cur = prev;
    if (completed) *completed = ndone;
    return AXON_OK;
}

extern "C" axon_status axon_server_result(axon_server* s, uint64_t id, uint32_t* spikes, uint64_t* row_digest,
                                          double* latency_s, uint32_t* T) {
    if (!s) return fail(AXON_ERR_INVALID_ARG, "null server");
    auto it = s->done.find(id);
    if (it == s->done.end()) return fail(AXON_ERR_NOT_READY, "request %llu not finished", (unsigned long long)id);
    const SrvResult& r = it->second;
    if (spikes) memcpy(spikes, r.spikes.data(), r.spikes.size() * 4);
    if (row_digest) memcpy(row_digest, r.digest.data(), r.digest.size() * 8);
    if (latency_s) *latency_s = r.done - r.arrival;
    if (T) *T = r.T;
    s->done.erase(it);
    return AXON_OK;
}

extern "C" axon_status axon_server_get_config(const axon_server* s, axon_server_config* cfg) {
    if (!s || !cfg) return fail(AXON_ERR_INVALID_ARG, "null argument");
    *cfg = s->cfg;
    return AXON_OK;
}

extern "C" axon_status axon_server_status(const axon_server* s, uint32_t* queued, uint32_t* inflight,
                                          uint64_t* reserved_bytes) {
    if (!s) return fail(AXON_ERR_INVALID_ARG, "null server");
    if (queued) *queued = s->queued;
    if (inflight) *inflight = s->inflight;
    if (reserved_bytes) *reserved_bytes = s->reserved;
    return AXON_OK;
}

// ================================================================== autotuning
// Probe-tile autotuner: runs the caller's probe batch once per tile size
// (AXON_PROBE_TILES sizes, kProbeTiles) and per occupancy of the propagation
// (kProbeWaves: waves per CU left resident), timing the device work with
// events, and reports the fastest pair.  Tiles of up to 8 probes list their work per
// (branch, word, probe) item; larger tiles walk the union of the probes'
// active branches and read each W3 line once for all of them.  Which wins
// depends on how much the probes' activity overlaps and on the number of
// probes, so it is measured, not assumed.  Each point is the median of
// kTuneReps timed runs after one untimed run (which also captures the
// step graph), and every result goes into the process's tuning table.

static constexpr int kTuneReps = 3;

// The tuning table: the best (tile, occupancy) found per (ncol, probes, T).
static std::mutex g_tune_mu;
static std::map<std::tuple<uint32_t, uint32_t, uint32_t>, axon_tune_entry> g_tune;

static void tune_record(const axon_tune_entry& e) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[std::make_tuple(e.ncol, e.probes, e.T)] = e;
}

// Times one (tile, occupancy) point on the first `count` stimuli: median and
// spread (max - min) of kTuneReps runs of T steps.
static axon_status tune_point(axon_net* n, uint32_t count, const uint32_t* stims, uint32_t T, uint32_t clamp_steps,
                              uint32_t use_bias, uint32_t tile, uint32_t waves, hipEvent_t e0, hipEvent_t e1,
                              double* med, double* spread) {
    axon_probes* b = nullptr;
    axon_status st = axon_probes_create(n, count, T, tile, &b);
    if (st != AXON_OK) return st;
    if ((st = axon_probes_set_occupancy(b, waves)) != AXON_OK) {
        axon_probes_destroy(b);
        return st;
    }
    std::vector<uint32_t> Ts(b->d.nslots, 0);
    std::fill(Ts.begin(), Ts.begin() + count, T);
    std::vector<uint32_t> pad((size_t)(b->d.nslots - count) * n->NCOL, 0);
    std::vector<double> t;
    for (int rep = 0; rep <= kTuneReps && st == AXON_OK; ++rep) {
        st = axon_probes_load(b, 0, count, stims, Ts.data(), clamp_steps, use_bias);
        if (st == AXON_OK && count < b->d.nslots)
            st = axon_probes_load(b, count, b->d.nslots - count, pad.data(), Ts.data() + count, clamp_steps, use_bias);
        if (st != AXON_OK) break;
        if (hipEventRecord(e0, n->stream) != hipSuccess) { st = fail(AXON_ERR_DEVICE, "event"); break; }
        if ((st = axon_probes_step(b, T)) != AXON_OK) break;
        if (hipEventRecord(e1, n->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            st = fail(AXON_ERR_DEVICE, "event");
            break;
        }
        float f = 0;
        (void)hipEventElapsedTime(&f, e0, e1);
        if (rep) t.push_back(f);  // run 0 captures the graph
    }
    axon_probes_destroy(b);
    if (st != AXON_OK) return st;
    std::sort(t.begin(), t.end());
    *med = t[t.size() / 2];
    *spread = t.back() - t.front();
    return AXON_OK;
}

// Every (tile, occupancy) point for the first `count` stimuli; the fastest
// goes into *best and the tuning table, ms[] (optional) gets every median.
static axon_status tune_batch(axon_net* n, uint32_t count, const uint32_t* stims, uint32_t T, uint32_t clamp_steps,
                              uint32_t use_bias, axon_tune_entry* best, double* ms) {
    HIPCHECK(hipSetDevice(n->device));
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(AXON_ERR_DEVICE, "hipEventCreate failed");
    }
    *best = axon_tune_entry{n->NCOL, count, T, 4u, 0u, (uint32_t)kTuneReps, 1e30, 0.0};
    axon_status st = AXON_OK;
    for (int c = 0; c < AXON_PROBE_TILES * AXON_PROBE_OCCUPANCIES && st == AXON_OK; ++c) {
        const int i = c / AXON_PROBE_OCCUPANCIES, o = c % AXON_PROBE_OCCUPANCIES;
        double med = 0, spread = 0;
        st = tune_point(n, count, stims, T, clamp_steps, use_bias, kProbeTiles[i], kProbeWaves[o], e0, e1, &med,
                        &spread);
        if (st != AXON_OK) break;
        if (ms) ms[c] = med;
        if (med < best->ms) {
            best->tile = kProbeTiles[i];
            best->waves_per_cu = kProbeWaves[o];
            best->ms = med;
            best->spread_ms = spread;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st == AXON_OK) tune_record(*best);
    return st;
}

static axon_status probe_autotune(axon_net* n, uint32_t count, const uint32_t* stims, uint32_t T,
                                  uint32_t clamp_steps, uint32_t use_bias, uint32_t* best_tile,
                                  uint32_t* best_waves_per_cu, double* ms) {
    if (!n || !stims || !best_tile || !best_waves_per_cu) return fail(AXON_ERR_INVALID_ARG, "null argument");
    if (count == 0) return fail(AXON_ERR_INVALID_ARG, "no probes");
    axon_tune_entry best{};
    axon_status st = tune_batch(n, count, stims, T, clamp_steps, use_bias, &best, ms);
    if (st != AXON_OK) return st;
    *best_tile = best.tile;
    *best_waves_per_cu = best.waves_per_cu;
    return AXON_OK;
}

// The batch-size axis: the same tuning for the first sizes[i] stimuli, each
// size its own table entry (a server's slots per network is such a size).
static axon_status probe_autotune_sizes(axon_net* n, uint32_t count, const uint32_t* stims, uint32_t T,
                                        uint32_t clamp_steps, uint32_t use_bias, const uint32_t* sizes,
                                        uint32_t nsizes, axon_tune_entry* out) {
    if (!n || !stims || !sizes || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    for (uint32_t i = 0; i < nsizes; ++i)
        if (sizes[i] == 0 || sizes[i] > count) return fail(AXON_ERR_INVALID_ARG, "size %u outside 1..%u", sizes[i], count);
    for (uint32_t i = 0; i < nsizes; ++i) {
        axon_status st = tune_batch(n, sizes[i], stims, T, clamp_steps, use_bias, &out[i], nullptr);
        if (st != AXON_OK) return st;
    }
    return AXON_OK;
}

// Nearest entry for ncol: the smallest distance in log2(probes) plus
// log2(T); an exact key has distance 0.
static bool tune_find(uint32_t ncol, uint32_t probes, uint32_t T, axon_tune_entry* out) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    double bestd = 1e30;
    for (const auto& kv : g_tune) {
        const axon_tune_entry& e = kv.second;
        if (e.ncol != ncol) continue;
        const double d = std::fabs(std::log2((double)e.probes) - std::log2((double)std::max(probes, 1u))) +
                         std::fabs(std::log2((double)e.T) - std::log2((double)std::max(T, 1u)));
        if (d < bestd) { bestd = d; *out = e; }
    }
    return bestd < 1e30;
}

extern "C" axon_status axon_tune_lookup(uint32_t ncol, uint32_t probes, uint32_t T, axon_tune_entry* out) {
    if (!out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    if (!tune_find(ncol, probes, T, out)) return fail(AXON_ERR_NOT_READY, "no tuning entry for ncol=%u", ncol);
    return AXON_OK;
}

extern "C" uint32_t axon_tune_entries(axon_tune_entry* out, uint32_t max) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    uint32_t i = 0;
    for (const auto& kv : g_tune) {
        if (out && i < max) out[i] = kv.second;
        ++i;
    }
    return i;
}

extern "C" void axon_tune_clear(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
}

// File form: a first line "axon-tune 1", then one line per entry:
// ncol probes T tile waves_per_cu reps ms spread_ms
static axon_status tune_save(const char* path) {
    if (!path) return fail(AXON_ERR_INVALID_ARG, "null path");
    FILE* f = fopen(path, "w");
    if (!f) return fail(AXON_ERR_IO, "cannot write %s", path);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    bool ok = fprintf(f, "axon-tune 1\n") > 0;
    for (const auto& kv : g_tune) {
        const axon_tune_entry& e = kv.second;
        ok = ok && fprintf(f, "%u %u %u %u %u %u %.6f %.6f\n", e.ncol, e.probes, e.T, e.tile, e.waves_per_cu, e.reps,
                           e.ms, e.spread_ms) > 0;
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? AXON_OK : fail(AXON_ERR_IO, "write to %s failed", path);
}

static bool valid_tile(uint32_t t) {
    for (uint32_t v : kProbeTiles)
        if (v == t) return true;
    return false;
}
static bool valid_waves(uint32_t w) {
    for (uint32_t v : kProbeWaves)
        if (v == w) return true;
    return false;
}

// Merges a saved table into this process's (a loaded key replaces the same
// key); nothing is merged unless every line parses and holds a valid choice.
static axon_status tune_load(const char* path) {
    if (!path) return fail(AXON_ERR_INVALID_ARG, "null path");
    FILE* f = fopen(path, "r");
    if (!f) return fail(AXON_ERR_IO, "cannot read %s", path);
    char magic[16] = {0};
    int ver = 0;
    std::vector<axon_tune_entry> got;
    bool ok = fscanf(f, "%15s %d", magic, &ver) == 2 && !strcmp(magic, "axon-tune") && ver == 1;
    while (ok) {
        axon_tune_entry e{};
        const int r = fscanf(f, "%u %u %u %u %u %u %lf %lf", &e.ncol, &e.probes, &e.T, &e.tile, &e.waves_per_cu,
                             &e.reps, &e.ms, &e.spread_ms);
        if (r == EOF) break;
        ok = r == 8 && e.ncol && e.probes && e.T && valid_tile(e.tile) && valid_waves(e.waves_per_cu);
        if (ok) got.push_back(e);
    }
    fclose(f);
    if (!ok) return fail(AXON_ERR_IO, "%s is not a tuning table", path);
    for (const axon_tune_entry& e : got) tune_record(e);
    return AXON_OK;
}

extern "C" axon_status axon_server_completed(const axon_server* s, uint64_t* ids, uint32_t max, uint32_t* n) {
    if (!s || !n || (max && !ids)) return fail(AXON_ERR_INVALID_ARG, "null argument");
    const uint32_t k = (uint32_t)std::min<size_t>(max, s->last_done.size());
    if (k) memcpy(ids, s->last_done.data(), (size_t)k * 8);
    *n = (uint32_t)s->last_done.size();
    return AXON_OK;
}

// ================================================================== snapshots
// Trained-network snapshot file, version 1 (layout in axon_api.h).  The planes
// are stored in the reference layout through axon_read_buffer, so a snapshot
// does not depend on the engine's internal tile order.
namespace {
constexpr char kSnapMagic[8] = {'A', 'X', 'O', 'N', 'S', 'N', 'A', 'P'};
constexpr uint32_t kSnapVersion = 1;
constexpr axon_buffer kSnapBufs[6] = {AXON_BUF_W0, AXON_BUF_W1, AXON_BUF_W2, AXON_BUF_W3, AXON_BUF_CONN, AXON_BUF_BIAS};

struct SnapHeader {
    char magic[8];
    uint32_t version, cfgBytes;
    axon_config cfg;
    uint64_t words[6];  // per stored buffer, kSnapBufs order
    uint32_t nstim, stimcols;
};

struct Fnv1a {  // 64-bit FNV-1a
    uint64_t h = 14695981039346656037ull;
    void add(const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    }
};

// buffer sizes (words) a network of this config has, kSnapBufs order
void snap_words(const axon_config& c, uint64_t w[6]) {
    const uint64_t N = (uint64_t)c.ncol * axon::NPC;
    for (int i = 0; i < 4; ++i) w[i] = (uint64_t)c.ncol * axon::NSLOT * axon::TILE_WORDS;
    w[4] = (uint64_t)c.ncol * axon::NSLOT;
    w[5] = N;
}
}  // namespace

extern "C" axon_status axon_net_config(const axon_net* n, axon_config* out) {
    if (!n || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    *out = n->cfg;
    return AXON_OK;
}

static axon_status net_save(axon_net* n, const char* path) {
    if (!n || !path) return fail(AXON_ERR_INVALID_ARG, "null argument");
    SnapHeader h{};
    memcpy(h.magic, kSnapMagic, 8);
    h.version = kSnapVersion;
    h.cfgBytes = (uint32_t)sizeof(axon_config);
    h.cfg = n->cfg;
    snap_words(n->cfg, h.words);
    h.nstim = (uint32_t)n->stimCols.size();
    h.stimcols = h.nstim ? (uint32_t)n->stimCols[0].size() : 0u;
    for (int i = 0; i < 6; ++i) {
        size_t w = 0;
        axon_status st = axon_buffer_words(n, kSnapBufs[i], &w);
        if (st != AXON_OK) return st;
        if (w < h.words[i]) return fail(AXON_ERR_INVALID_ARG, "buffer %d holds %zu words, expected %llu", i, w,
                                        (unsigned long long)h.words[i]);
    }
    // written under a temporary name and renamed when complete, so a reader
    // never sees a partial snapshot at `path`
    const std::string tmp = std::string(path) + ".partial";
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return fail(AXON_ERR_IO, "cannot create %s: %s", tmp.c_str(), strerror(errno));
    Fnv1a sum;
    bool ok = true;
    auto put = [&](const void* p, size_t bytes) {
        if (ok && fwrite(p, 1, bytes, f) != bytes) ok = false;
        sum.add(p, bytes);
    };
    put(&h, sizeof h);
    std::vector<uint32_t> buf;
    axon_status st = AXON_OK;
    for (int i = 0; i < 6 && st == AXON_OK && ok; ++i) {
        buf.resize(h.words[i]);
        st = axon_read_buffer(n, kSnapBufs[i], buf.data(), buf.size() * 4);
        if (st == AXON_OK) put(buf.data(), buf.size() * 4);
    }
    for (const std::vector<uint32_t>& s : n->stimCols) put(s.data(), s.size() * 4);
    const uint64_t digest = sum.h;
    if (ok && fwrite(&digest, 1, 8, f) != 8) ok = false;
    if (fclose(f) != 0) ok = false;
    if (st != AXON_OK || !ok) {
        remove(tmp.c_str());
        return st != AXON_OK ? st : fail(AXON_ERR_IO, "writing %s failed", tmp.c_str());
    }
    if (rename(tmp.c_str(), path) != 0) {
        const int err = errno;
        remove(tmp.c_str());
        return fail(AXON_ERR_IO, "cannot rename %s to %s: %s", tmp.c_str(), path, strerror(err));
    }
    return AXON_OK;
}

static axon_status net_load(const char* path, axon_net** out) {
    if (!path || !out) return fail(AXON_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    FILE* f = fopen(path, "rb");
    if (!f) return fail(AXON_ERR_IO, "cannot open %s: %s", path, strerror(errno));
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, fclose);
    SnapHeader h{};
    if (fread(&h, 1, sizeof h, f) != sizeof h) return fail(AXON_ERR_BAD_SNAPSHOT, "%s: shorter than a header", path);
    if (memcmp(h.magic, kSnapMagic, 8) != 0) return fail(AXON_ERR_BAD_SNAPSHOT, "%s: not an AXON snapshot", path);
    if (h.version != kSnapVersion)
        return fail(AXON_ERR_BAD_SNAPSHOT, "%s: snapshot version %u, this engine reads %u", path, h.version,
                    kSnapVersion);
    if (h.cfgBytes != sizeof(axon_config)) return fail(AXON_ERR_BAD_SNAPSHOT, "%s: config size differs", path);
    if (axon_config_validate(&h.cfg) != AXON_OK)
        return fail(AXON_ERR_BAD_SNAPSHOT, "%s: stored config is invalid: %s", path, g_err.c_str());
    uint64_t want[6];
    snap_words(h.cfg, want);
    for (int i = 0; i < 6; ++i)
        if (h.words[i] != want[i]) return fail(AXON_ERR_BAD_SNAPSHOT, "%s: buffer %d size differs from its config", path, i);
    if (h.nstim != h.cfg.stim || h.stimcols != h.cfg.stimcols)
        return fail(AXON_ERR_BAD_SNAPSHOT, "%s: stimulus table size differs from its config", path);
    // the file must hold exactly header + buffers + stimulus table + checksum
    // before anything is sized from the header
    uint64_t expect = sizeof h + 8u + 4ull * h.nstim * h.stimcols;
    for (int i = 0; i < 6; ++i) expect += 4ull * h.words[i];
    if (fseeko(f, 0, SEEK_END) != 0) return fail(AXON_ERR_IO, "%s: cannot seek", path);
    const off_t flen = ftello(f);
    if (flen < 0 || (uint64_t)flen != expect)
        return fail(AXON_ERR_BAD_SNAPSHOT, "%s: %lld bytes, its header describes %llu", path, (long long)flen,
                    (unsigned long long)expect);
    if (fseeko(f, (off_t)sizeof h, SEEK_SET) != 0) return fail(AXON_ERR_IO, "%s: cannot seek", path);
    Fnv1a sum;
    sum.add(&h, sizeof h);
    std::vector<std::vector<uint32_t>> bufs(6);
    for (int i = 0; i < 6; ++i) {
        bufs[i].resize(h.words[i]);
        if (fread(bufs[i].data(), 4, bufs[i].size(), f) != bufs[i].size())
            return fail(AXON_ERR_BAD_SNAPSHOT, "%s: truncated", path);
        sum.add(bufs[i].data(), bufs[i].size() * 4);
    }
    std::vector<uint32_t> stim((size_t)h.nstim * h.stimcols);
    if (fread(stim.data(), 4, stim.size(), f) != stim.size()) return fail(AXON_ERR_BAD_SNAPSHOT, "%s: truncated", path);
    sum.add(stim.data(), stim.size() * 4);
    uint64_t digest = 0;
    if (fread(&digest, 1, 8, f) != 8) return fail(AXON_ERR_BAD_SNAPSHOT, "%s: truncated", path);
    if (digest != sum.h) return fail(AXON_ERR_BAD_SNAPSHOT, "%s: checksum mismatch", path);
    if (fgetc(f) != EOF) return fail(AXON_ERR_BAD_SNAPSHOT, "%s: trailing bytes after the checksum", path);
    axon_net* n = nullptr;
    axon_status st = axon_net_create(&h.cfg, &n);
    if (st != AXON_OK) return st;
    for (size_t q = 0; q < n->stimCols.size(); ++q)
        if (!std::equal(n->stimCols[q].begin(), n->stimCols[q].end(), stim.begin() + q * h.stimcols)) {
            axon_net_destroy(n);
            return fail(AXON_ERR_BAD_SNAPSHOT, "%s: stimulus columns of pattern %zu differ from its config's", path, q);
        }
    for (int i = 0; i < 6 && st == AXON_OK; ++i) st = axon_write_buffer(n, kSnapBufs[i], bufs[i].data(), bufs[i].size() * 4);
    if (st != AXON_OK) {
        axon_net_destroy(n);
        return st;
    }
    *out = n;
    return AXON_OK;
}

// ================================================================== memory probes
// Propagation probe: the counterpart of the reference's probe05 (dense and
// activity-swept binary propagation, bench/probe05_prop.mm:41-70, 200-212)
// and probe06 (the same spike count scattered over neurons or gathered into
// whole columns, bench/probe06_columnar.mm:117-133, 180-200), measured on this
// engine's own propagation kernel and synapse layout rather than on a
// separate benchmark kernel: `spikes` becomes the presynaptic activity of a
// step (the spike words and the per-column "spiked" bytes the kernel stages),
// and the network's propagation runs on it.
static axon_status prop_probe(axon_net* n, const uint32_t* spikes, uint32_t reps, double* us, uint64_t* lines) {
    if (!n || !spikes || !us || !lines || reps == 0u) return fail(AXON_ERR_INVALID_ARG, "null argument or reps 0");
    const uint32_t NC = n->NCOL;
    HIPCHECK(hipSetDevice(n->device));
    HIPCHECK(hipStreamSynchronize(n->stream));
    std::vector<uint8_t> nz(NC);
    for (uint32_t c = 0; c < NC; ++c) nz[c] = spikes[c] != 0u ? 1u : 0u;
    HIPCHECK(put(n, n->d.spk[0], spikes, (size_t)NC * 4));
    HIPCHECK(put(n, n->d.colNZ, nz.data(), NC));
    HIPCHECK(zero(n, n->d.cntProp, (size_t)NC * 8));
    Params P = n->P;
    P.step = 0;
    HIPCHECK(axon::launch_drive(n->d, P, 0, n->stream));  // untimed; its line count is the answer
    HIPCHECK(hipStreamSynchronize(n->stream));
    uint64_t pr[2];
    axon_status st = counter_pairs(n, n->d.cntProp, pr, 0);
    if (st != AXON_OK) return st;
    *lines = pr[1];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHECK(hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, n->stream);
    for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) e = axon::launch_drive(n->d, P, 0, n->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, n->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(AXON_ERR_DEVICE, "propagation probe: %s", hipGetErrorString(e));
    *us = 1e3 * (double)ms / reps;
    HIPCHECK(zero(n, n->d.cntProp, (size_t)NC * 8));
    return AXON_OK;
}

extern "C" axon_status axon_mem_probe(uint32_t kind, uint64_t bytes, uint32_t reps, double* gbps) {
    if (!gbps || kind > 2u || reps == 0u) return fail(AXON_ERR_INVALID_ARG, "kind 0..2, reps >= 1, gbps not null");
    if (bytes < 1024u || bytes % 1024u || bytes > (64ull << 30))
        return fail(AXON_ERR_INVALID_ARG, "bytes must be a multiple of 1 KiB in 1 KiB..64 GiB");
    int dev = 0;
    axon_status st = select_device(&dev);
    if (st != AXON_OK) return st;
    void* p = nullptr;
    uint32_t* out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (s) (void)hipStreamSynchronize(s);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
        if (out) (void)hipFree(out);
        if (p) (void)hipFree(p);
    };
    if (hipMalloc(&p, bytes) != hipSuccess || hipMalloc((void**)&out, 64) != hipSuccess) {
        cleanup();
        return fail(AXON_ERR_OUT_OF_MEMORY, "memory probe: hipMalloc(%llu)", (unsigned long long)bytes);
    }
    const uint32_t blocks = kind == 0 ? 4096u : 8192u, iters = 64u;
    double best = 0.0;
    hipError_t e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(p, 0, bytes, s);
    if (e == hipSuccess) e = axon::launch_mem_probe(kind, (const uint32_t*)p, bytes, blocks, iters, out, s);  // warm-up
    for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) {
        e = hipEventRecord(e0, s);
        if (e == hipSuccess) e = axon::launch_mem_probe(kind, (const uint32_t*)p, bytes, blocks, iters, out, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && ms > 0.f)
            best = std::max(best, (double)axon::mem_probe_bytes(kind, bytes, blocks, iters) / (ms * 1e-3) / 1e9);
    }
    cleanup();
    if (e != hipSuccess) return fail(AXON_ERR_DEVICE, "memory probe: %s", hipGetErrorString(e));
    *gbps = best;
    return AXON_OK;
}

extern "C" axon_status axon_sync_probe(uint32_t kind, uint32_t groups, uint32_t workgroups, uint32_t reps,
                                       double* us) {
    if (!us || kind > 3u || reps == 0u || groups == 0u || groups > 64u || workgroups == 0u || workgroups > 65536u)
        return fail(AXON_ERR_INVALID_ARG, "kind 0..3, groups 1..64, workgroups 1..65536, reps >= 1, us not null");
    int dev = 0;
    axon_status st = select_device(&dev);
    if (st != AXON_OK) return st;
    uint32_t *out = nullptr, *cnt = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (s) (void)hipStreamSynchronize(s);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
        if (cnt) (void)hipFree(cnt);
        if (out) (void)hipFree(out);
    };
    const size_t outWords = (size_t)groups * workgroups * 264u;
    if (hipMalloc((void**)&out, outWords * 4) != hipSuccess || hipMalloc((void**)&cnt, (size_t)groups * 128) != hipSuccess) {
        cleanup();
        return fail(AXON_ERR_OUT_OF_MEMORY, "synchronisation probe: hipMalloc");
    }
    hipError_t e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, (size_t)groups * 128, s);
    if (e == hipSuccess) e = axon::launch_sync_probe(kind, out, cnt, groups, workgroups, s);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) e = axon::launch_sync_probe(kind, out, cnt, groups, workgroups, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    cleanup();
    if (e != hipSuccess) return fail(AXON_ERR_DEVICE, "synchronisation probe: %s", hipGetErrorString(e));
    *us = 1e3 * (double)ms / reps;
    return AXON_OK;
}

// ================================================================== guarded entry points
// The entry points whose host allocations depend on their arguments or on a
// file: the bodies above, behind guarded() (no exception crosses the C ABI).
extern "C" axon_status axon_net_create(const axon_config* c, axon_net** out) {
    return guarded("axon_net_create", [&] { return net_create(c, out); });
}
extern "C" axon_status axon_net_save(axon_net* n, const char* path) {
    return guarded("axon_net_save", [&] { return net_save(n, path); });
}
extern "C" axon_status axon_net_load(const char* path, axon_net** out) {
    return guarded("axon_net_load", [&] { return net_load(path, out); });
}
extern "C" axon_status axon_wdiag(axon_net* n, const uint32_t* memb, const uint32_t* asm_spk, uint32_t nstim,
                                  uint32_t sel, uint32_t hist[48]) {
    return guarded("axon_wdiag", [&] { return wdiag(n, memb, asm_spk, nstim, sel, hist); });
}
extern "C" axon_status axon_probes_create(axon_net* n, uint32_t slots, uint32_t max_t, uint32_t tile,
                                          axon_probes** out) {
    return guarded("axon_probes_create", [&] { return probes_create(n, slots, max_t, tile, out); });
}
extern "C" axon_status axon_probes_run(axon_probes* b, uint32_t count, const uint32_t* stims, uint32_t T,
                                       uint32_t clamp_steps, uint32_t use_bias, uint32_t* hist, uint32_t* spikes,
                                       double* ms) {
    return guarded("axon_probes_run",
                   [&] { return probes_run(b, count, stims, T, clamp_steps, use_bias, hist, spikes, ms); });
}
extern "C" axon_status axon_multi_create(axon_net* const* nets, uint32_t nnet, axon_multi** out) {
    return guarded("axon_multi_create", [&] { return multi_create(nets, nnet, out); });
}
extern "C" axon_status axon_multi_train(axon_multi* b, const uint32_t* epochs, uint32_t extra, double* ms) {
    return guarded("axon_multi_train", [&] { return multi_train(b, epochs, extra, ms); });
}
extern "C" axon_status axon_server_create(axon_net* const* nets, uint32_t nnet, const axon_server_config* cfg,
                                          axon_server** out) {
    return guarded("axon_server_create", [&] { return server_create(nets, nnet, cfg, out); });
}
extern "C" axon_status axon_server_submit(axon_server* s, uint32_t net, const uint32_t* stim, uint32_t T,
                                          uint32_t clamp_steps, uint32_t use_bias, double arrival_s, uint64_t* id) {
    return guarded("axon_server_submit",
                   [&] { return server_submit(s, net, stim, T, clamp_steps, use_bias, arrival_s, id); });
}
extern "C" axon_status axon_server_tick(axon_server* s, uint32_t* completed) {
    return guarded("axon_server_tick", [&] { return server_tick(s, completed); });
}
extern "C" axon_status axon_probe_autotune(axon_net* n, uint32_t count, const uint32_t* stims, uint32_t T,
                                           uint32_t clamp_steps, uint32_t use_bias, uint32_t* best_tile,
                                           uint32_t* best_waves_per_cu,
                                           double ms[AXON_PROBE_TILES * AXON_PROBE_OCCUPANCIES]) {
    return guarded("axon_probe_autotune", [&] {
        return probe_autotune(n, count, stims, T, clamp_steps, use_bias, best_tile, best_waves_per_cu, ms);
    });
}
extern "C" axon_status axon_probe_autotune_sizes(axon_net* n, uint32_t count, const uint32_t* stims, uint32_t T,
                                                 uint32_t clamp_steps, uint32_t use_bias, const uint32_t* sizes,
                                                 uint32_t nsizes, axon_tune_entry* out) {
    return guarded("axon_probe_autotune_sizes", [&] {
        return probe_autotune_sizes(n, count, stims, T, clamp_steps, use_bias, sizes, nsizes, out);
    });
}
extern "C" axon_status axon_tune_save(const char* path) {
    return guarded("axon_tune_save", [&] { return tune_save(path); });
}
extern "C" axon_status axon_tune_load(const char* path) {
    return guarded("axon_tune_load", [&] { return tune_load(path); });
}
extern "C" axon_status axon_prop_probe(axon_net* n, const uint32_t* spikes, uint32_t reps, double* us,
                                       uint64_t* lines) {
    return guarded("axon_prop_probe", [&] { return prop_probe(n, spikes, reps, us, lines); });
}
