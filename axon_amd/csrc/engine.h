// Internal interface between the C-ABI host layer (api.cpp) and the HIP
// kernels (kernels.hip).  Not installed; the public surface is axon_api.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace axon {

constexpr uint32_t NPC = 32;       // neurons per column, mm:32
constexpr uint32_t NSLOT = 32;     // dendritic branches per neuron, mm:34
constexpr uint32_t PRESYN = 256;   // presynaptic neurons per branch, mm:36
constexpr uint32_t FANIN = 8192;   // mm:37
constexpr uint32_t COLS_PER_GRP = 8;
// A (column, branch) tile is 32 neurons x 256 presynaptic bits per plane,
// stored word-major: plane[(tile * 8 + w) * 32 + n] holds the 32 bits of
// presynaptic column 8g + w (g = the branch's group) for neuron n, so one
// presynaptic column's word for all 32 neurons is one 128-byte line, and a
// kernel reads only the lines of presynaptic columns that spiked.
constexpr uint32_t TILE_WORDS = 256;  // uint32 per tile per plane

// Plasticity work lists per step parity, and the columns one list can hold
// (column c is in list (c / 4) % NLISTS).
constexpr uint32_t NLISTS = 8;
__host__ __device__ constexpr uint32_t list_cols(uint32_t ncol) { return (ncol + 31u) / 32u * 4u; }

// Per-step kernel parameters.  Same 26 words as the reference Params
// (mm:56-83); the fields no kernel reads (learn, freeze, theta) are kept so
// the struct stays a faithful mirror.
struct Params {
    uint32_t k, theta, dthr, boost, mode, inCols, step, learn, depress, targetCols, clamp,
        freeze, biasUp, biasDown, biasShift, histOn, useBias, phase, bandLo, bandHi,
        marginGate, hetero, ltdRep, biasDownEvery, facil, pad;
};

// Device pointers of one network instance.
struct NetDev {
    uint32_t* W0; uint32_t* W1; uint32_t* W2; uint32_t* W3;  // word-major tiles
    uint32_t* conn;
    uint32_t* bias;
    uint32_t* spk[2];
    uint32_t* acc;
    uint32_t* bbits;
    uint32_t* drive;
    uint32_t* theta;
    uint32_t* marg;
    uint32_t* stim;
    uint32_t* stats;
    uint32_t* hist;
    uint32_t* driveLog;
    uint32_t* refBall;
    uint32_t* mult;
    // Work counters (no reference counterpart), accumulated across steps:
    // per column, [2c] (column, branch) tiles propagation evaluated and
    // [2c + 1] 128-byte W3 lines it read; the same pair for the tiles and word
    // lines (4 planes each) plasticity rewrote.
    uint32_t* cntProp;   // ncol * 2
    uint32_t* cntLearn;  // ncol * 2
    uint8_t* colNZ;      // ncol (+pad to 16): column spiked in the last step
    // Per (column, branch) tile: the 8-bit mask of the branch group's columns
    // that spiked in the step propagation last read (written by propagation,
    // read by plasticity), and whether every counter of the tile is known to
    // lie in [bandLo, bandHi] (set by a full-tile plasticity pass with the
    // band clip, cleared by host writes of the planes).
    uint8_t* firedW;     // ncol * 32
    uint8_t* tileBand;   // ncol * 32
    // Plasticity work of step t with side statistics: the learning columns, in
    // list t & 1.  Each parity holds 8 lists (column c in list (c / 4) % 8) so
    // that columns take list slots on 8 different counters.  (Without side
    // statistics plasticity needs no list: the next propagation rewrites its
    // own columns' voted tiles.)
    uint32_t* learnList; // [2][8][list_cols(ncol)]
    uint32_t* learnCnt;  // [2][8] list lengths, 32 words apart (reset by the step after next's propagation)
    // The part of the plasticity work done at a presentation's last timestep
    // (in its emit launch): per column (tiles, lines); the work counters add
    // it in.
    uint32_t* cntLast;   // ncol * 2
    uint32_t ncol, incols, inbr, topkpar;
    // active: batched training only -- this round's pattern of the network
    // (~0 = idle), written by k_mb_begin into the batch's device copy of the
    // descriptors so that a kernel learns it with the descriptor itself.
    uint32_t nstats, ndlog, maxt, active;
    // 1: `acc` holds 16-bit accumulators (u16[ncol*32]), chosen by the host
    // when no accumulator can reach 2^16; 0: 32-bit
    uint32_t acc16, pad[3];
};

// Device pointers of a probe batch: `nslots` independent readout
// presentations against one network (its W3, conn and bias).
struct ProbeDev {
    const uint32_t* W3;
    const uint32_t* conn;
    const uint32_t* bias;
    uint32_t* spk;       // [slot][ncol]   spikes of the last completed step
    uint32_t* gact;      // [slot][gw]     1 bit per column that spiked in the last step
    uint32_t* acc;       // [slot][ncol*32]
    uint32_t* drive;     // [slot][ncol]
    uint32_t* stim;      // [slot][ncol]
    uint32_t* hist;      // [slot][maxt][ncol]
    uint32_t* theta;     // [slot][2]
    uint32_t* st_t;      // [slot] steps done
    uint32_t* st_T;      // [slot] steps to run
    uint32_t* st_clamp;  // [slot] clamped steps
    uint32_t* st_bias;   // [slot] use the intrinsic bias (0/1)
    uint32_t* selDone;   // [slot] workgroups of a split select that finished this step (reset by the last)
    uint32_t nslots, ncol, incols, maxt, gw;
    uint32_t k, facil, dthr, boost, mode, biasShift, target, topkpar;
    // 1: `acc` holds 16-bit accumulators ([slot][ncol*32] uint16), chosen by
    // the host when no accumulator can reach 2^16 (probe_acc16), which halves
    // the propagation's largest write; 0: 32-bit
    uint32_t acc16, pad;
};

// Lock-step batch of independent networks (same dynamics parameters).
struct MultiDev {
    const NetDev* nets;       // [nnet], device copy
    const uint32_t* sched;    // [rounds][nnet]: pattern of this round, ~0 = idle
    uint32_t* round;          // device round counter
    const uint64_t* stimTabs; // [nnet]: device address of the [stim][ncol] clamp patterns
    uint32_t nnet, pad;
};

// ---- launchers (kernels.hip) ; all asynchronous on `s` ----
hipError_t launch_multi_begin(const MultiDev& m, uint32_t ncol, uint32_t nstats, hipStream_t s);
// PL (optional): the parameters of step t - 1, whose plasticity the launch
// applies first (deferred plasticity); the last step's runs with its emit
// (launch_multi_emit_learn).  P0 (optional, t = 1): step 0's parameters, a
// clamped step whose work the launch takes (launch_multi_emit_first's).
hipError_t launch_multi_drive(const MultiDev& m, const Params& P, uint32_t t, uint32_t ncol, hipStream_t s,
                              const Params* PL = nullptr, const Params* P0 = nullptr);
hipError_t launch_multi_topk(const MultiDev& m, const Params& P, hipStream_t s);
hipError_t launch_multi_emit(const MultiDev& m, const Params& P, uint32_t t, bool learn, uint32_t ncol,
                             hipStream_t s);
// the last step's plasticity of a round, which also advances the round counter (as launch_multi_end)
// the last step of a round when it learns: emit and plasticity in one launch, which also advances the
// round counter
hipError_t launch_multi_emit_learn(const MultiDev& m, const Params& P, uint32_t t, uint32_t ncol, hipStream_t s);
hipError_t launch_multi_end(const MultiDev& m, hipStream_t s);
// timestep 0 of a round together with its begin work (k_mb_emit_first); then
// launch_multi_begin and step 0's other launches are not needed
hipError_t launch_multi_emit_first(const MultiDev& m, const Params& P, bool learn, uint32_t ncol, hipStream_t s);
// advances slots [0, nact) (nact <= nslots; the others must be idle);
// wavesPerCU (0: no limit) caps the propagation's resident waves per CU
hipError_t launch_probe_step(const ProbeDev& d, uint32_t tile, uint32_t nact, uint32_t wavesPerCU, hipStream_t s);
hipError_t launch_probe_reset(const ProbeDev& d, uint32_t first, uint32_t count, hipStream_t s);
// Allocation-free probe runs: the static per-GPU scratch array (words), its
// device address on `dev` (the current device; looked up once per device),
// the start-of-run slot setup, and a word copy that is a kernel (no memset or
// runtime copy node when a caller captures the stream).
constexpr size_t NOALLOC_SCRATCH_WORDS = (size_t)(32u << 20) / 4u;  // 32 MiB
hipError_t noalloc_scratch(int dev, uint32_t** out);
hipError_t launch_probe_setup(const ProbeDev& d, uint32_t count, uint32_t T, uint32_t clampSteps, uint32_t useBias,
                              hipStream_t s);
hipError_t launch_copy_words(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t s);
hipError_t launch_probe_admit(const ProbeDev& d, const uint32_t* list, uint32_t n, hipStream_t s);
hipError_t launch_probe_gather(const ProbeDev& d, const uint32_t* slots, const uint32_t* offs, uint32_t n,
                               uint32_t* out, hipStream_t s);
hipError_t launch_init(const NetDev& d, uint32_t nand, uint32_t seed, hipStream_t s);
hipError_t launch_clear(const NetDev& d, hipStream_t s);
// e0 / e1 (optional): timing events recorded at the start / end of the kernel.
// PL (optional): the parameters of step t - 1, whose plasticity (without side
// statistics) the propagation applies first to its columns' voted tiles.
// P0 (optional, t = 1, no PL): the parameters of step 0, clamped to stim0
// (d.stim when null, else a pattern-table row, which the launch copies into
// d.stim), which lies inside the input columns; the launch also does step
// 0's work, which then has no launch.
hipError_t launch_drive(const NetDev& d, const Params& P, uint32_t t, hipStream_t s, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr, const Params* PL = nullptr, const Params* P0 = nullptr,
                        const uint32_t* stim0 = nullptr);
// first: step 0 of a presentation without side statistics, which needs no
// clear, propagation or plasticity launch (k_emit_first)
hipError_t launch_emit(const NetDev& d, const Params& P, uint32_t t, bool learn, bool stats, hipStream_t s,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, bool first = false);
// Plasticity of step t with side statistics, over the learning-column lists
// (k_learn); without them it is deferred into the next propagation
// (launch_drive's PL) or, at the last step, runs with the emit
// (launch_emit_learn).
hipError_t launch_learn(const NetDev& d, const Params& P, uint32_t t, hipStream_t s, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr);
// The last timestep of a presentation without side statistics, when it
// learns: emit and that step's plasticity in one launch (k_emit_learn)
hipError_t launch_emit_learn(const NetDev& d, const Params& P, uint32_t t, hipStream_t s, hipEvent_t e0 = nullptr,
                             hipEvent_t e1 = nullptr);

hipError_t launch_wstat(const NetDev& d, uint32_t* out, hipStream_t s);
hipError_t launch_within(const NetDev& d, const uint32_t* ball, const uint32_t* postMask, uint32_t useMask,
                         uint32_t* out, hipStream_t s);
hipError_t launch_indeg(const NetDev& d, uint32_t* out, hipStream_t s);
hipError_t launch_untouched(const NetDev& d, uint32_t* out, hipStream_t s);
// anyAsm: ncol words of scratch (the OR of the assemblies' ballots per column)
hipError_t launch_wdiag(const NetDev& d, const uint32_t* memb, const uint32_t* asmSpk, uint32_t nstim,
                        uint32_t sel, uint32_t* anyAsm, uint32_t* hist, hipStream_t s);
hipError_t launch_fill(uint32_t* p, uint32_t value, size_t nwords, hipStream_t s);
// memory probes: kind 0 streaming read, 1 128-byte line gathers, 2 1 KiB tile gathers
hipError_t launch_mem_probe(uint32_t kind, const uint32_t* p, size_t bytes, uint32_t blocks, uint32_t iters,
                            uint32_t* out, hipStream_t s);
size_t mem_probe_bytes(uint32_t kind, size_t bytes, uint32_t blocks, uint32_t iters);
// synchronisation probe: out holds groups * nwg * 264 words, cnt groups * 32 words
hipError_t launch_sync_probe(uint32_t kind, uint32_t* out, uint32_t* cnt, uint32_t groups, uint32_t nwg,
                             hipStream_t s);
hipError_t launch_checksum(const uint32_t* p, size_t n, unsigned long long* out, hipStream_t s);

hipError_t launch_bstest(uint32_t op, uint32_t lo, uint32_t hi, uint32_t* a0, uint32_t* a1, uint32_t* a2,
                         uint32_t* a3, const uint32_t* m, size_t nwords, hipStream_t s);
hipError_t launch_kwta_test(const uint32_t* v, uint32_t ncols, uint32_t k, uint32_t* ball, uint32_t* vk,
                            uint32_t* vk1, hipStream_t s);
hipError_t launch_topk_test(const uint32_t* drive, uint32_t ninst, uint32_t ncol, uint32_t incols,
                            uint32_t target, uint32_t par, uint32_t* theta, hipStream_t s);

}  // namespace axon
