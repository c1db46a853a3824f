/*
 * axon_api.h -- C ABI of the MI355X AXON engine (libaxon.so).
 *
 * Plain C: fixed-width integers, plain pointers and sizes, opaque handles, no
 * C++ or torch types.  Every entry point returns an axon_status and never
 * aborts; axon_last_error() gives the message for the calling thread.
 *
 * Each declaration names the reference interface it replaces.  The reference
 * is a single Objective-C++/Metal program (src/axon.mm): its "interface" for
 * the hot path is the key=value CLI, the present() lambda and the kernel
 * buffer-slot tables.  Citations are src/axon.mm line numbers (mm:N).
 *
 * Ownership: the engine owns all device memory.  Output pointers are caller
 * host buffers sized from the config (NCOL = ncol, N = ncol * 32); NULL means
 * "do not copy this output".
 */
#ifndef AXON_API_H
#define AXON_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AXON_ABI_VERSION 1u

typedef enum axon_status {
    AXON_OK = 0,
    AXON_ERR_INVALID_CONFIG = 1, /* a config the reference would run into undefined behaviour */
    AXON_ERR_OUT_OF_MEMORY = 2,
    AXON_ERR_BUDGET_EXCEEDED = 3, /* serving admission: request larger than the HBM budget */
    AXON_ERR_DEVICE = 4,          /* HIP runtime error or no usable gfx950 device */
    AXON_ERR_INVALID_ARG = 5,
    AXON_ERR_CAPACITY = 6, /* e.g. T beyond the history capacity max_t */
    AXON_ERR_NOT_READY = 7, /* serving: no result available yet */
    AXON_ERR_IO = 8,        /* a file could not be opened, read or written */
    AXON_ERR_BAD_SNAPSHOT = 9 /* not a snapshot of this version, wrong sizes, or a checksum mismatch */
} axon_status;

/* Every key=value tunable of the reference CLI, mm:687-735 (defaults mm:688-728,
 * derived rules mm:731-735).  Field order matches the oracle's ora_config. */
typedef struct axon_config {
    uint32_t ncol, incols, nand, target, k, dthr, boost, mode;
    uint32_t epochs, T, clampsteps, stim, seed, quiet, split, unionchar;
    uint32_t biasup, biasdown, biasshift, biasdownevery, inbr, stimcols, diag, profile;
    uint32_t topkpar, vtopk, contrast, bandlo, bandhi, margingate, hetero, vbs;
    uint32_t ltdrep, ltdevery, ktest, facil, deep, wave, wavet, trainclamp;
    uint32_t max_t; /* history capacity in steps; the reference fixes MAXT = 64 (mm:766) */
    uint32_t reserved[7];
} axon_config;

typedef struct axon_net axon_net;

uint32_t axon_abi_version(void);
const char* axon_status_str(axon_status s);
const char* axon_last_error(void);

/* ---- configuration: replaces argu() (mm:660-667) and the tunables block (mm:687-735) ---- */
void axon_config_defaults(axon_config* c);
/* key=value, strncmp prefix + '=', last occurrence wins, atoi -> u32, then the
 * derived rules of mm:731-735 (trainclamp default, biasup=1 auto). */
axon_status axon_config_parse(int argc, const char* const* argv, axon_config* c);
axon_status axon_config_validate(const axon_config* c);

/* ---- the whole program: replaces the reference's main() (mm:669-1911) ----
 * Parses the key=value arguments (argv[0] is the program name, as for the
 * CLI), builds and trains the network, runs the reference's protocol and
 * prints its report (every block the keys switch on, byte for byte the
 * reference's lines) to report_path (NULL = stdout).  save=<path> and
 * load=<path> store or reuse the network after its training epochs.
 * trained: NULL, or a network created with the same training keys whose
 * epochs are done (e.g. by axon_multi_train with extra = 0); the run uses it
 * in place of training and continues on it (its extra learning presentation,
 * mm:1033, changes it); the caller keeps ownership.  Keys that only select
 * readouts or printing (quiet, split, unionchar, diag, profile, vtopk, vbs,
 * deep, wave, wavet) may differ from the trained network's. */
axon_status axon_run(int argc, const char* const* argv, axon_net* trained, const char* report_path);

/* ---- network: replaces the buffer block mm:749-768, k_init_w/k_init_conn
 * dispatch mm:770-782 and the stimulus draw mm:793-799 ---- */
axon_status axon_net_create(const axon_config* c, axon_net** out);
void axon_net_destroy(axon_net* n);
/* the network's configuration (as created or loaded) */
axon_status axon_net_config(const axon_net* n, axon_config* out);

/* Trained-network snapshots.  No counterpart in the reference, which retrains
 * from its seed on every run; for serving residency and sweeps.  Version-1
 * file, little-endian: a header (magic "AXONSNAP", version, the network's
 * axon_config, the buffer sizes), the four planes in the reference layout
 * (mm:170), conn, bias, the stimulus columns of every pattern (mm:793-799),
 * then a 64-bit FNV-1a checksum of every byte before it.
 * axon_net_load checks magic, version, sizes, the config and the checksum
 * before any device work, then creates the network from the stored config
 * and writes its buffers; the stimulus columns it regenerates from the config
 * must equal the stored ones. */
axon_status axon_net_save(axon_net* n, const char* path);
axon_status axon_net_load(const char* path, axon_net** out);

/* ---- one presentation: replaces present() mm:815-888 ----
 * use_bias: 0, 1 or 2 (= auto: on while learning, off at readout, mm:822).
 * stim: optional NCOL-word clamp pattern replacing loadStim(pattern, keep, seed)
 * (mm:800-804); NULL uses the reference draw. */
typedef struct axon_present_args {
    uint32_t pattern;
    uint32_t T;
    uint32_t clamp_steps;
    uint32_t learn;
    uint32_t freeze; /* accepted for signature parity; no kernel reads it (mm:840) */
    uint32_t seed;
    uint32_t use_bias;
    uint32_t flags; /* AXON_PRESENT_ASYNC: enqueue only, outputs must be NULL */
    double keep;
    const uint32_t* stim;
} axon_present_args;

#define AXON_PRESENT_ASYNC 1u

typedef struct axon_present_out {
    uint32_t* spikes;    /* NCOL: final ballots, (T&1)?B:A (mm:891) */
    uint32_t* hist;      /* T*NCOL: per-step ballots (k_emit hist, mm:224) */
    uint32_t* stats;     /* stats words (>= 1024, mm:214-233, 491-494) */
    uint32_t* drive_log; /* driveLog words (>= 256, mm:216-220) */
    uint32_t* theta;     /* 2 words (mm:291/351) */
    double ms_per_step;  /* host wall time per step, as present() returns */
} axon_present_out;

axon_status axon_present(axon_net* n, const axon_present_args* a, axon_present_out* out);
axon_status axon_sync(axon_net* n);

/* host writes into Shared buffers of the reference (bRefBall mm:1700, bMult
 * mm:1604/1656, bDriveLog memset mm:1713) */
axon_status axon_set_ref_ball(axon_net* n, const uint32_t* ball /* NCOL */);
axon_status axon_set_mult(axon_net* n, const uint32_t* mult /* N */);
axon_status axon_clear_drive_log(axon_net* n);

/* raw engine buffers, for snapshots and parity checks */
typedef enum axon_buffer {
    AXON_BUF_W0 = 0, AXON_BUF_W1, AXON_BUF_W2, AXON_BUF_W3, AXON_BUF_CONN, AXON_BUF_SPK_A,
    AXON_BUF_SPK_B, AXON_BUF_DRIVE, AXON_BUF_BBITS, AXON_BUF_ACC, AXON_BUF_BIAS, AXON_BUF_MARG,
    AXON_BUF_REFBALL, AXON_BUF_DRIVELOG, AXON_BUF_MULT, AXON_BUF_STIM, AXON_BUF_STATS,
    AXON_BUF_HIST, AXON_BUF_THETA, AXON_BUF_COUNT
} axon_buffer;
axon_status axon_buffer_words(const axon_net* n, axon_buffer which, size_t* nwords);
axon_status axon_read_buffer(axon_net* n, axon_buffer which, void* dst, size_t bytes);
axon_status axon_write_buffer(axon_net* n, axon_buffer which, const void* src, size_t bytes);
/* switches the column top-K between the parallel (1) and serial (0) selection
 * semantics, the pTopK switch of mm:806-809 used by the vtopk check mm:966-992 */
axon_status axon_net_set_topk_parallel(axon_net* n, uint32_t par);
/* stimCols[p] of mm:798 (stimcols entries) */
axon_status axon_stim_cols(const axon_net* n, uint32_t p, uint32_t* out);

/* ---- measurement (no reference counterpart; the reference timed whole
 * presentations with steady_clock, mm:834/886, and per-kernel Metal
 * timestamps under profile=1, mm:1162-1340) ---- */
/* out[0]: (column, branch) tiles read by propagation, out[1]: tiles rewritten
 * by plasticity, accumulated since creation or the last reset.  Each tile is
 * 8192 synapses (32 neurons x 256 presynaptic bits). */
axon_status axon_work_counters(axon_net* n, uint64_t out[2], uint32_t reset);
/* No reference counterpart: the HBM lines behind those tiles.  out[0]: 128-byte
 * W3 lines propagation read (one per spiking presynaptic column of an active
 * tile), out[1]: word lines plasticity rewrote (each 4 planes x 128 bytes read
 * and written).  Reset clears these and the work counters together. */
axon_status axon_line_counters(axon_net* n, uint64_t out[2], uint32_t reset);
/* No reference counterpart: the part of the plasticity counts above done at
 * the last timestep of presentations without side statistics, inside that
 * timestep's emit launch (out[0] tiles, out[1] word lines); the other
 * timesteps' plasticity runs inside the next timestep's propagation launch.
 * Cleared by the resets above. */
axon_status axon_last_step_counters(axon_net* n, uint64_t out[2]);
/* enable = 1: presentations run eagerly with HIP events around every kernel
 * and wait for the device before returning (an AXON_PRESENT_ASYNC
 * presentation still skips its side statistics and outputs);
 * axon_kernel_times returns summed device ms and launch counts for
 * [0] propagation (without side statistics including the previous
 * timestep's plasticity, and at the second timestep the first's work when
 * that runs inside it), [1] the first timestep of a presentation without
 * side statistics as a launch of its own (in place of the clear, the
 * propagation and the plasticity of that step; none when the first timestep
 * is clamped to a stimulus inside the input columns: then the second
 * timestep's propagation does its work), [2] column top-K + emit (k-WTA, bias),
 * [3] plasticity launches: with side statistics the plasticity of each
 * learning timestep; without, the last timestep's emit launch, which also
 * runs that timestep's plasticity (its emit is not in [2]). */
axon_status axon_net_set_profiling(axon_net* n, uint32_t enable);
axon_status axon_kernel_times(axon_net* n, double ms[4], uint64_t launches[4], uint32_t reset);

/* ---- readout kernels ---- */
axon_status axon_wstat(axon_net* n, uint32_t* functional);  /* k_wstat mm:641-646, wstat() mm:784-791 */
/* k_within mm:524-562: post_mask NULL means useMask = 0 (the deep call, mm:1702-1709) */
axon_status axon_within(axon_net* n, const uint32_t* ball, const uint32_t* post_mask, uint32_t out[16]);
axon_status axon_indeg(axon_net* n, uint32_t* out /* N */);     /* k_indeg mm:566-581 */
axon_status axon_untouched(axon_net* n, uint32_t* out /* N */); /* k_untouched mm:501-519 */
/* k_wdiag mm:591-638: memb N words, asm_spk nstim*NCOL words, hist 48 words */
axon_status axon_wdiag(axon_net* n, const uint32_t* memb, const uint32_t* asm_spk, uint32_t nstim,
                       uint32_t sel, uint32_t hist[48]);

/* ---- probe batches: many readout presentations against one network ----
 * Batching extension with no reference counterpart; each probe is exactly the
 * reference's readout present(q, keep, T, clampSteps, learn=false, ...)
 * (mm:815-888 with learn = 0) on its own clamp pattern, and its spike
 * ballots are bit-identical to that single presentation.  Probes are packed in
 * tiles of `tile` (8/16/32/64) slots; one W3 tile read serves a whole probe
 * tile.  Slots advance together one timestep per step; a slot stops at its own
 * T, so probes of different lengths share launches.  The margin, stats and
 * drive-log side outputs of a presentation are not produced for probes. */
typedef struct axon_probes axon_probes;
axon_status axon_probes_create(axon_net* n, uint32_t slots, uint32_t max_t, uint32_t tile, axon_probes** out);
void axon_probes_destroy(axon_probes* b);
axon_status axon_probes_info(const axon_probes* b, uint32_t* slots, uint32_t* max_t, uint32_t* tile);
/* stims: count*NCOL clamp patterns; T: count step counts (<= max_t); use_bias 0/1 */
axon_status axon_probes_load(axon_probes* b, uint32_t first, uint32_t count, const uint32_t* stims,
                             const uint32_t* T, uint32_t clamp_steps, uint32_t use_bias);
axon_status axon_probes_step(axon_probes* b, uint32_t steps); /* asynchronous */
axon_status axon_probes_sync(axon_probes* b);
/* spikes: count*NCOL final ballots; hist: count*max_t*NCOL; steps_done: count */
axon_status axon_probes_read(axon_probes* b, uint32_t first, uint32_t count, uint32_t* spikes, uint32_t* hist,
                             uint32_t* steps_done);
/* load `count` probes with the same T into slots 0.., run them to completion;
 * hist is count*T*NCOL (packed), ms the device wall time of the steps */
axon_status axon_probes_run(axon_probes* b, uint32_t count, const uint32_t* stims, uint32_t T,
                            uint32_t clamp_steps, uint32_t use_bias, uint32_t* hist, uint32_t* spikes, double* ms);
/* Allocation-free probe run.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same results as
 * axon_probes_run on a batch of `tile` probes per propagation wave: `count`
 * readout presentations of T steps against n.  Every piece of probe state
 * lives in one static device buffer of AXON_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends); a run whose state does not fit returns
 * AXON_ERR_CAPACITY (per probe slot: 3*NCOL + N + NCOL/32 words, plus T*NCOL
 * when d_hist is NULL; slots are count rounded up to the tile).
 * Device pointers: d_stims count*NCOL clamp patterns (input); d_hist
 * count*T*NCOL words (each probe's T rows, packed) or NULL; d_spikes
 * count*NCOL final ballots or NULL.  stream: a hipStream_t, NULL = the
 * network's stream.  Asynchronous: enqueues kernels only (no allocation, no
 * memset, no host synchronisation), so it may be captured.  All calls on one
 * GPU share the buffer, which keeps the last call's contents: calls that
 * overlap on one GPU (different streams) race; serialise them. */
#define AXON_NOALLOC_SCRATCH_BYTES (32u << 20)
axon_status axon_probes_run_noalloc(axon_net* n, uint32_t count, const uint32_t* d_stims, uint32_t T,
                                    uint32_t clamp_steps, uint32_t use_bias, uint32_t tile, uint32_t* d_hist,
                                    uint32_t* d_spikes, void* stream);
/* waves per CU the batch's propagation may keep resident: 0 (as many as
 * fit), 32 or 16; fewer leaves each wave more of the CU's caches */
axon_status axon_probes_set_occupancy(axon_probes* b, uint32_t waves_per_cu);
/* autotuner: runs this probe batch with every tile size (1/2/4/8/16/32/64
 * probes per propagation wave) and every occupancy (0/32/16 waves per CU),
 * times the device work (median of 3 timed runs after an untimed one),
 * returns the fastest pair and records it in the tuning table under (ncol,
 * count, T); ms[tile_index * AXON_PROBE_OCCUPANCIES + occupancy_index] in
 * those orders */
#define AXON_PROBE_TILES 7
#define AXON_PROBE_OCCUPANCIES 3
axon_status axon_probe_autotune(axon_net* n, uint32_t count, const uint32_t* stims, uint32_t T,
                                uint32_t clamp_steps, uint32_t use_bias, uint32_t* best_tile,
                                uint32_t* best_waves_per_cu, double ms[AXON_PROBE_TILES * AXON_PROBE_OCCUPANCIES]);
/* Tuning table: the autotuners record, per (ncol, probes, T), the fastest
 * (tile, waves per CU) with the median time of its timed runs (reps of them)
 * and their spread (max - min).  The table lives in the process; save and
 * load keep it across processes (text, first line "axon-tune 1").  A server
 * created with tile = 0 takes the entry nearest to (ncol, slots_per_net,
 * max_t).  Extension with no reference counterpart. */
typedef struct axon_tune_entry {
    uint32_t ncol, probes, T, tile, waves_per_cu, reps;
    double ms, spread_ms;
} axon_tune_entry;
/* the batch-size axis: axon_probe_autotune for the first sizes[i] stimuli
 * (1 <= sizes[i] <= count), one entry each, written to out[i] */
axon_status axon_probe_autotune_sizes(axon_net* n, uint32_t count, const uint32_t* stims, uint32_t T,
                                      uint32_t clamp_steps, uint32_t use_bias, const uint32_t* sizes, uint32_t nsizes,
                                      axon_tune_entry* out);
/* the entry for ncol nearest to (probes, T) in log2(probes) + log2(T);
 * AXON_ERR_NOT_READY when the table has none for ncol */
axon_status axon_tune_lookup(uint32_t ncol, uint32_t probes, uint32_t T, axon_tune_entry* out);
/* copies up to max entries, returns how many the table holds */
uint32_t axon_tune_entries(axon_tune_entry* out, uint32_t max);
void axon_tune_clear(void);
axon_status axon_tune_save(const char* path);
/* merges a saved table (a key loaded replaces the same key); a file with any
 * line that does not parse or names an unknown tile/occupancy merges nothing */
axon_status axon_tune_load(const char* path);
/* recall cue: pattern's stimulus columns with `nreplace` of them replaced by
 * other input columns (mt19937(seed) draws); pattern 0xFFFFFFFF = a random cue
 * of stimcols input columns.  Extends loadStim (mm:800-804). */
axon_status axon_probe_cue(const axon_net* n, uint32_t pattern, uint32_t nreplace, uint32_t seed, uint32_t* stim);

/* ---- many networks trained together ----
 * Batching extension with no reference counterpart: independent networks
 * (different seeds and numbers of stored patterns, identical dynamics
 * parameters) advance in the same launches.  Each network follows exactly the
 * reference's training protocol -- epochs x stim learning presentations
 * (pattern k mod stim, keep 1.0, mm:1009-1010) and, with extra = 1, the
 * additional presentation of pattern 0 (mm:1033) -- so its state is
 * bit-identical to training it alone.  Uneven lengths are handled by
 * per-network work queues (a device schedule table). */
typedef struct axon_multi axon_multi;
axon_status axon_multi_create(axon_net* const* nets, uint32_t nnet, axon_multi** out);
void axon_multi_destroy(axon_multi* b);
axon_status axon_multi_train(axon_multi* b, const uint32_t* epochs /* nnet */, uint32_t extra, double* ms);
/* order-independent 64-bit checksum of an engine buffer (state comparison) */
axon_status axon_checksum(axon_net* n, axon_buffer which, uint64_t* sum);

/* Read bandwidth of the engine's access shapes over a `bytes` buffer (a
 * multiple of 1 KiB, at most 64 GiB): kind 0 a streaming pass, 1 128-byte
 * line gathers, 2 1 KiB tile gathers (pseudo-random places).  Best of `reps`
 * timed launches after one warm-up, in GB/s (1e9 bytes/s).  The counterpart
 * of the reference's Apple probes (bench/probe04_roofline.mm): calibration
 * of the percent-of-peak figures, not part of the hot path. */
axon_status axon_mem_probe(uint32_t kind, uint64_t bytes, uint32_t reps, double* gbps);

/* Propagation probe, the counterpart of the reference's probe05 (dense and
 * activity-swept propagation, bench/probe05_prop.mm:41-70) and probe06
 * (scattered vs columnar activity of one spike count,
 * bench/probe06_columnar.mm:117-133), run on this engine's propagation
 * kernel: `spikes` (NCOL words) is the presynaptic activity of a step; one
 * untimed launch gives *lines, the 128-byte W3 lines a launch reads, then
 * `reps` launches are timed back to back, *us = mean microseconds per launch.
 * Overwrites the network's step state (spikes, drive, accumulators, work
 * counters); the planes, conn and bias are unchanged. */
axon_status axon_prop_probe(axon_net* n, const uint32_t* spikes, uint32_t reps, double* us, uint64_t* lines);

/* No reference counterpart: the cost of cross-workgroup synchronisation at a
 * propagation launch's shape (groups x workgroups of 256 threads, each
 * storing what a batched propagation workgroup stores), the measurement
 * behind running a timestep as separate launches instead of one persistent
 * kernel with grid-wide waits.  kind 0: the stores alone; 1: plus one
 * arrival atomic per workgroup on its group's counter; 2: plus an
 * agent-scope release fence before it; 3: plus a full agent-scope fence.
 * No workgroup waits for another.  *us = mean microseconds per launch over
 * `reps` back-to-back launches after one warm-up. */
axon_status axon_sync_probe(uint32_t kind, uint32_t groups, uint32_t workgroups, uint32_t reps, double* us);

/* ---- serving: continuous batching of recall queries ----
 * Extension with no reference counterpart.  Trained networks stay resident;
 * each query is one readout presentation (a probe) of T steps on one network.
 * axon_server_tick admits queued queries in arrival order into free slots
 * while the sum of running footprints stays within hbm_budget (0 = unlimited),
 * advances every running query by steps_per_tick steps (networks on separate
 * streams, so their work overlaps), gathers the queries those steps finish,
 * and collects the results the previous tick gathered: ticks are pipelined
 * two deep, so a query is returned by the tick after the one that finished
 * it, and the host never waits for the tick it has just enqueued.  `inflight`
 * (axon_server_status) counts queries admitted and not yet returned; tick
 * until it and `queued` are both 0 to drain the server.  A result is the final
 * spike ballots plus one 64-bit checksum per step's ballot row (the
 * axon_checksum formula over that row), bit-identical to the same query run
 * alone.  Latency is completion time minus the arrival time given at submit
 * (seconds on axon_server_clock; negative = now). */
typedef struct axon_server axon_server;
typedef struct axon_server_config {
    uint32_t slots_per_net; /* probe slots per network */
    uint32_t max_t;         /* longest query, steps */
    uint32_t tile;          /* probe tile: 1, 2, 4, 8, 16, 32 or 64; 0 = the tuning table's choice */
    uint32_t steps_per_tick;
    uint64_t hbm_budget;    /* bytes of in-flight query state, 0 = no limit */
    uint32_t waves_per_cu;  /* probe propagation occupancy: 0 (no limit), 32 or 16 (set by the table when tile = 0) */
    uint32_t reserved[3];
} axon_server_config;
axon_status axon_server_create(axon_net* const* nets, uint32_t nnet, const axon_server_config* cfg,
                               axon_server** out);
void axon_server_destroy(axon_server* s);
double axon_server_clock(const axon_server* s);
axon_status axon_server_submit(axon_server* s, uint32_t net, const uint32_t* stim, uint32_t T, uint32_t clamp_steps,
                               uint32_t use_bias, double arrival_s, uint64_t* id);
axon_status axon_server_tick(axon_server* s, uint32_t* completed);
/* ids of the queries the last tick returned (up to max; *n = how many) */
axon_status axon_server_completed(const axon_server* s, uint64_t* ids, uint32_t max, uint32_t* n);
/* spikes NCOL words, row_digest T words; AXON_ERR_NOT_READY until finished */
axon_status axon_server_result(axon_server* s, uint64_t id, uint32_t* spikes, uint64_t* row_digest,
                               double* latency_s, uint32_t* T);
/* the configuration in use (tile and waves_per_cu as chosen when tile was 0) */
axon_status axon_server_get_config(const axon_server* s, axon_server_config* cfg);
axon_status axon_server_status(const axon_server* s, uint32_t* queued, uint32_t* inflight,
                               uint64_t* reserved_bytes);

/* ---- kernel test harnesses (replace the vbs/vtopk self-checks mm:903-994) ---- */
/* k_bstest mm:424-434: op 0 inc, 1 dec, 2 clip[lo,hi]; nwords u32 per plane, in place */
axon_status axon_bs_test(uint32_t op, uint32_t lo, uint32_t hi, uint32_t* a0, uint32_t* a1, uint32_t* a2,
                         uint32_t* a3, const uint32_t* m, size_t nwords);
/* kwta_ballot_margin mm:134-144 on ncols groups of 32 values */
axon_status axon_kwta_test(const uint32_t* v, uint32_t ncols, uint32_t k, uint32_t* ball, uint32_t* vk,
                           uint32_t* vk1);
/* k_topk_par / k_topk_serial mm:259-352 on ninst drive vectors of ncol words */
axon_status axon_topk_test(const uint32_t* drive, uint32_t ninst, uint32_t ncol, uint32_t incols,
                           uint32_t target, uint32_t par, uint32_t* theta /* 2*ninst */);

#ifdef __cplusplus
}
#endif
#endif
