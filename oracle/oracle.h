/*
 * AXON CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded C++ restatement of the reference engine's update rules
 * (src/axon.mm, Metal kernels mm:45-647 and host protocol mm:660-1906), with an
 * optional column-parallel variant that is bit-identical to it.  It exists to
 * check the HIP engine and to serve as the timed CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product (axon_amd/) never links or calls it.
 *
 * Parity pins (what fixes each part independently of the restatement itself):
 *   hash32 / init      lowbias32 closed form; counter values in {5,6,9,14};
 *                      untrained within-assembly density 6.3613 % (paper:253-255)
 *   mt19937            standard known answer: 10000th output of seed 5489 is 4123659995
 *   libc++ shuffle     reproduced end-to-end by the golden row (README.md:57-66)
 *   bs_inc/dec/clip    closed forms min(15,v+1), max(0,v-1), clamp (mm:945-947)
 *   top-K theta        nth_element closed form; serial == parallel (mm:958-994)
 *   k-of-32            brute-force rank and popcount(ballot) == k (probe05:235-250)
 *   whole pipeline     golden row 82.0905 / 0.746 / 0.042 / 0.661 / 118 and
 *                      paper Table 1 rows (paper:259-272)
 */
#ifndef AXON_ORACLE_H
#define AXON_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Every key=value tunable of the reference CLI (mm:687-735), in one struct.
 * The field order is identical to axon_config in include/axon_api.h. */
typedef struct ora_config {
    uint32_t ncol, incols, nand, target, k, dthr, boost, mode;
    uint32_t epochs, T, clampsteps, stim, seed, quiet, split, unionchar;
    uint32_t biasup, biasdown, biasshift, biasdownevery, inbr, stimcols, diag, profile;
    uint32_t topkpar, vtopk, contrast, bandlo, bandhi, margingate, hetero, vbs;
    uint32_t ltdrep, ltdevery, ktest, facil, deep, wave, wavet, trainclamp;
    uint32_t max_t; /* history capacity in steps (reference: MAXT = 64) */
    uint32_t reserved[7];
} ora_config;

typedef struct ora_net ora_net;

/* argu() semantics: strncmp prefix + '=', last occurrence wins, atoi -> u32,
 * followed by the derived rules of mm:731-735. */
void ora_config_defaults(ora_config* c);
void ora_config_parse(int argc, const char* const* argv, ora_config* c);
/* 0 when the config is one the reference can run without undefined behaviour. */
int ora_config_check(const ora_config* c, char* msg, size_t msg_len);

/* nthreads <= 1: single-threaded transcription; > 1: column-parallel. */
ora_net* ora_net_create(const ora_config* c, int nthreads);
void ora_net_destroy(ora_net* n);
void ora_net_set_threads(ora_net* n, int nthreads);

/* present() of mm:815-888.  use_bias: 0, 1 or 2 (=auto).  stim_override, when
 * non-NULL, replaces loadStim (NCOL words).  Returns the number of steps run. */
uint32_t ora_present(ora_net* n, uint32_t pattern, double keep, uint32_t T,
                     uint32_t clamp_steps, uint32_t learn, uint32_t seed,
                     uint32_t use_bias, const uint32_t* stim_override);

/* Host views of the engine state (all little-endian u32 words). */
enum {
    ORA_W0 = 0, ORA_W1, ORA_W2, ORA_W3, ORA_CONN, ORA_SPK_A, ORA_SPK_B, ORA_DRIVE,
    ORA_BBITS, ORA_ACC, ORA_BIAS, ORA_MARG, ORA_REFBALL, ORA_DRIVELOG, ORA_MULT,
    ORA_STIM, ORA_STATS, ORA_HIST, ORA_THETA, ORA_NBUF
};
uint32_t* ora_buffer(ora_net* n, int which, size_t* nwords);
const uint32_t* ora_final_spikes(ora_net* n, uint32_t T); /* (T&1)?B:A, mm:891 */
void ora_stim_cols(ora_net* n, uint32_t p, uint32_t* out); /* stimcols entries */

/* Work counters (no reference counterpart): tiles read by propagation and
 * tiles rewritten by plasticity since creation / the last reset. */
void ora_work_counters(ora_net* n, uint64_t out[2], int reset);

/* Readout kernels (mm:501-646).  post_mask == NULL means useMask = 0. */
void ora_within(ora_net* n, const uint32_t* ball, const uint32_t* post_mask, uint32_t out[16]);
uint32_t ora_wstat(ora_net* n);
void ora_indeg(ora_net* n, uint32_t* out);
void ora_untouched(ora_net* n, uint32_t* out);
void ora_wdiag(ora_net* n, const uint32_t* memb, const uint32_t* asm_spk, uint32_t nstim,
               uint32_t sel, uint32_t hist[48]);

/* bit-serial counter ops on word arrays (k_bstest, mm:424-434): op 0 inc, 1 dec, 2 clip */
void ora_bs_apply(uint32_t op, uint32_t lo, uint32_t hi, uint32_t* a0, uint32_t* a1,
                  uint32_t* a2, uint32_t* a3, const uint32_t* m, size_t nwords);
/* top-K theta over drive[incols..ncol) with parallel (par=1) or serial semantics */
void ora_topk(const uint32_t* drive, uint32_t ncol, uint32_t incols, uint32_t target,
              int par, uint32_t theta[2]);
/* k-of-32 ballot with margin (mm:134-144) */
uint32_t ora_kwta(const uint32_t v[32], uint32_t k, uint32_t* vk, uint32_t* vk1);

/* RNG restatements: mt19937 word stream, libc++ shuffle, uniform_real(0,1). */
uint32_t ora_mt_nth(uint32_t seed, uint32_t n); /* n-th output (1-based) */
void ora_libcxx_shuffle(uint32_t* a, uint32_t n, uint32_t mt_seed);
double ora_uniform01(uint32_t seed, uint32_t draw_index);

/* The reference protocol (mm:996-1858) reduced to the state it mutates plus the
 * readouts the paper pins.  Writes the report into buf (NUL-terminated). */
typedef struct ora_pinned {
    double within_pct, between_pct, fin, pin, fout, pout;
    double recall75, other75, margin75;   /* A4 completion row, keep 0.75 */
    double drive_t[40]; uint32_t drive_tt;  /* deep drive trace */
    double wave_ov[60]; uint32_t wave_tw;   /* WAVE ov(clampRef) per step */
    double sustained;                       /* printed(ov[last]) / printed(ov[clamp]) */
    double asm_jacc;                        /* A3 measured mean pairwise Jaccard */
    uint32_t wstat_final;
} ora_pinned;
int ora_run_protocol(const ora_config* c, int nthreads, ora_pinned* out, char* buf, size_t buf_len);

#ifdef __cplusplus
}
#endif
#endif
