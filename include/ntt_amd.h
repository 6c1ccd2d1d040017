/*
 * ntt_amd.h — C ABI of the MI355X (gfx950) RNS multi-limb negacyclic NTT/INTT.
 *
 * Plain C types only (pointers, sizes, integers).  Device pointers are HIP
 * device addresses; `stream` is a hipStream_t passed as void* (NULL = the
 * legacy default stream, which is what the reference launches on).
 *
 * Data layout (same as the reference): operand A[limb * N + i], uint32,
 * N = 2^logN, transformed in place.  Forward: natural-order input,
 * bit-reversed-order output.  Inverse: bit-reversed input, natural output.
 *   Montgomery variant (src/v04_montgomery.cu): forward input normal domain,
 *     output Montgomery form (x * 2^32 mod q); inverse input Montgomery form,
 *     output normal domain after the N^-1 scaling.
 *   Plain variant (src/v03_template_config.cu): normal domain throughout.
 * Inputs are expected in [0, q); outputs are always canonical, in [0, q).
 *
 * Errors: every entry point returns an ntt_status.  Launch entry points are
 * asynchronous on `stream`; a kernel fault surfaces at the caller's next
 * synchronisation, as with the reference (src/v04_montgomery.cu:990-991).
 * Where the reference calls std::exit (src/v04_montgomery.cu:685-688,
 * 702-706, 787-790) this library returns a status instead.
 */
#ifndef NTT_AMD_H
#define NTT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NTT_OK = 0,
    NTT_ERR_INVALID_ARGUMENT = 1,   /* null pointer, logN out of range, zero limbs */
    NTT_ERR_MODULUS_CONGRUENCE = 2, /* q != 1 mod 2N            (v04:685-688) */
    NTT_ERR_PRIMITIVE_ROOT = 3,     /* root does not give a primitive 2N-th root (v04:702-706) */
    NTT_ERR_MODULUS_RANGE = 4,      /* q even, q <= 1 or q >= 2^31 (Montgomery bound, v04:110-118) */
    NTT_ERR_CONFIG = 5,             /* phase config inconsistent with logN or the device */
    NTT_ERR_HIP = 6,                /* a HIP runtime call failed */
    NTT_ERR_OUT_OF_MEMORY = 7,
    NTT_ERR_BUDGET = 8,             /* request larger than the server's HBM budget */
    NTT_ERR_NOT_FOUND = 9,          /* unknown request id */
    NTT_ERR_BUSY = 10,              /* queue full / try again */
    NTT_ERR_NO_DEVICE = 11,
    NTT_ERR_SCRATCH_LIMIT = 12      /* more jobs or limbs than ntt_batch_run_static's fixed tables hold */
} ntt_status;

typedef enum { NTT_DOMAIN_PLAIN = 0, NTT_DOMAIN_MONTGOMERY = 1 } ntt_domain;
typedef enum { NTT_FORWARD = 0, NTT_INVERSE = 1 } ntt_direction;

/* Kernel family a plan launches on.  TWO_PASS: the reference's two-phase
 * structure (column pass + row pass, one HBM round trip each).  FUSED: one
 * persistent workgroup per CU holds 2^16 words (one or more whole limbs) in
 * registers and LDS, one HBM round trip per transform (10 <= logN <= 16);
 * at logN 17 a workgroup transforms one limb as two 2^16-word halves in
 * turn, 1.5 round trips (the intermediate half goes through the output).
 * AUTO: the autotuned choice for the launch size, else FUSED where supported. */
typedef enum { NTT_ENGINE_AUTO = 0, NTT_ENGINE_TWO_PASS = 1, NTT_ENGINE_FUSED = 2 } ntt_engine;

/* Where the fused engine takes twiddles in the round whose threads each need
 * their own (the upper levels of the table).  TABLE: from the table.  OT: on
 * the fly, tw[i] = mont_mul(tw[i mod 64], msb[i / 64]), from the table's
 * first 64 words and an N/64-word appendix msb[h] = tw[64 h] that the
 * context stores after every table of 2^10 words or more.  AUTO (the
 * default): the tuned choice for the launch size, else OT. */
typedef enum { NTT_TWIDDLE_AUTO = -1, NTT_TWIDDLE_TABLE = 0, NTT_TWIDDLE_OT = 1 } ntt_twiddle_mode;

/* ModulusConfig, src/ntt_config.h:6-9 */
typedef struct {
    uint32_t qi;
    uint32_t primitive_root;
} ntt_modulus_config;

/* PhaseConfig, src/ntt_config.h:12-16 (all log2).  On gfx950:
 *   radix_stages   stages of the transform done by this phase;
 *   stage_merging  stages done per register round (2^S words per thread);
 *   warp_batching  column phase: log2 of consecutive columns per tile
 *                  (6 = one wave64 covers 64 columns, 256 B per row);
 *                  row phase: log2 of rows per workgroup.
 * Workgroup = 2^(radix_stages - stage_merging + warp_batching) threads,
 * LDS = workgroup << stage_merging words, as in src/v04_montgomery.cu:582-601. */
typedef struct {
    int32_t radix_stages;
    int32_t stage_merging;
    int32_t warp_batching;
} ntt_phase_config;

/* TransformConfig, src/ntt_config.h:18-23 */
typedef struct {
    ntt_phase_config ntt_phase1;   /* forward column phase (outer stages) */
    ntt_phase_config ntt_phase2;   /* forward row phase (inner stages)    */
    ntt_phase_config intt_phase1;  /* inverse row phase (inner stages)    */
    ntt_phase_config intt_phase2;  /* inverse column phase + N^-1 scaling */
} ntt_transform_config;

/* ModulusRuntime, src/v04_montgomery.cu:17-26 */
typedef struct {
    uint32_t qi, primitive_root, q_inv, r2, psi, psi_inv, omega, degree_inv;
} ntt_modulus_runtime;

/* ---------------------------------------------------------------------------
 * Library / device information
 * ------------------------------------------------------------------------- */
const char* ntt_status_string(int status);
/* Text of the last HIP error this thread's calls turned into NTT_ERR_HIP. */
const char* ntt_last_hip_error(void);
const char* ntt_version(void);
/* Max logN supported (20).  The fused engine covers 10..17 (17: one
 * workgroup per limb, transformed as two 2^16-word halves in turn); the two-pass
 * engine every size, with the reference's phase split up to logN 18 and the
 * wave64 default split (ntt_default_transform_config) up to 20. */
int ntt_max_log_n(void);

/* ---------------------------------------------------------------------------
 * Configuration API — src/ntt_config.h
 * ------------------------------------------------------------------------- */
/* default_moduli(), src/ntt_config.h:75-94.  Writes up to `cap` entries and
 * returns the number available (16). */
int ntt_default_moduli(ntt_modulus_config* out, int cap);

/* DefaultNTTLaunchConfig<LogN>, src/ntt_config.h:62-73, with warp_batching
 * re-derived for wave64 (column phases 6, row phases 0). */
int ntt_default_transform_config(int logN, ntt_transform_config* out);

/* The reference's literal DefaultNTTLaunchConfig<LogN> values (W = 4 / 0). */
int ntt_reference_transform_config(int logN, ntt_transform_config* out);

/* Checks what the reference leaves unchecked (radix1 + radix2 == logN,
 * stage_merging <= radix_stages, 2^S <= 32, workgroup <= 1024, LDS fits,
 * column tile <= columns).  Returns NTT_OK or NTT_ERR_CONFIG. */
int ntt_validate_transform_config(int logN, const ntt_transform_config* cfg);

/* The `count` largest primes q < bound with q = 1 (mod 2^(logN+1)), in
 * descending order, each with its smallest primitive root.  Returns the number
 * written (fewer if the range runs out). */
int ntt_generate_moduli(int logN, uint32_t bound, int count, ntt_modulus_config* out);

/* ---------------------------------------------------------------------------
 * Host precompute — src/v04_montgomery.cu:125-141, 660-740
 * ------------------------------------------------------------------------- */
int ntt_make_modulus_runtime(int logN, ntt_modulus_config cfg, ntt_modulus_runtime* out);
/* tw[i] = psi^brv(i, logN) mod q, normal domain (v04:720-727). */
int ntt_make_twiddle_table(int logN, uint32_t q, uint32_t psi, uint32_t* tw_out);
/* Limb-major Montgomery-form tables for `limbs` moduli, as the reference's
 * main() builds them (v04:893-915).  Any output pointer may be NULL. */
int ntt_make_device_tables_host(int logN, const ntt_modulus_config* moduli, size_t limbs,
                                int domain, uint32_t* twiddles, uint32_t* inv_twiddles,
                                uint32_t* moduli_out, uint32_t* q_inv, uint32_t* r2,
                                uint32_t* degree_inv);
uint32_t ntt_mont_q_inv(uint32_t q);           /* v04:125-131 */
uint32_t ntt_to_mont(uint32_t x, uint32_t q);  /* v04:133-136 */
uint32_t ntt_mont_r2(uint32_t q);              /* v04:138-141 */
uint32_t ntt_from_mont(uint32_t x, uint32_t q, uint32_t q_inv); /* v04:143-146 */

/* ---------------------------------------------------------------------------
 * Drop-in launchers.  Same argument meaning as the reference's; logN and the
 * phase config (NULL = ntt_default_transform_config) replace the template
 * parameter LaunchConfig.
 * ------------------------------------------------------------------------- */
/* launch_ntt<LC>, src/v04_montgomery.cu:603-626 (Montgomery-form twiddles,
 * fused x R^2 input conversion). */
int ntt_launch_ntt_montgomery(uint32_t* A_dev, const uint32_t* twiddle_dev,
                              const uint32_t* moduli_dev, const uint32_t* q_inv_dev,
                              const uint32_t* r2_dev, size_t limb_count, int logN,
                              const ntt_transform_config* cfg, void* stream);
/* launch_intt<LC>, src/v04_montgomery.cu:628-658 (degree_inv_dev holds
 * to_mont(N^-1); NormalizeKernel is fused into the last phase). */
int ntt_launch_intt_montgomery(uint32_t* A_dev, const uint32_t* inv_twiddle_dev,
                               const uint32_t* moduli_dev, const uint32_t* q_inv_dev,
                               const uint32_t* degree_inv_dev, size_t limb_count, int logN,
                               const ntt_transform_config* cfg, void* stream);
/* launch_ntt<LC>, src/v03_template_config.cu:549-571 (normal-domain twiddles). */
int ntt_launch_ntt_plain(uint32_t* A_dev, const uint32_t* twiddle_dev,
                         const uint32_t* moduli_dev, size_t limb_count, int logN,
                         const ntt_transform_config* cfg, void* stream);
/* launch_intt<LC>, src/v03_template_config.cu:573-603 (degree_inv_dev normal domain). */
int ntt_launch_intt_plain(uint32_t* A_dev, const uint32_t* inv_twiddle_dev,
                          const uint32_t* moduli_dev, const uint32_t* degree_inv_dev,
                          size_t limb_count, int logN, const ntt_transform_config* cfg,
                          void* stream);

/* ---------------------------------------------------------------------------
 * Context + plans: device-resident Montgomery twiddle cache keyed by
 * (q, 2N-th root, logN, direction) (two primitive roots of one q give two
 * tables), replacing the per-call table setup in the reference's main()
 * (src/v04_montgomery.cu:882-916, 947-967).
 * ------------------------------------------------------------------------- */
typedef struct ntt_context ntt_context;
typedef struct ntt_plan ntt_plan;

int ntt_context_create(int device, ntt_context** out);
int ntt_context_destroy(ntt_context* ctx);
/* Bytes of device memory held by the twiddle cache. */
size_t ntt_context_twiddle_bytes(const ntt_context* ctx);
/* Number of distinct (q, logN) tables in the cache. */
size_t ntt_context_twiddle_entries(const ntt_context* ctx);

/* cfg NULL = the autotuned config for (logN, limbs) if one is loaded, else
 * the default.  Validates every modulus (the checks the reference exits on). */
int ntt_plan_create(ntt_context* ctx, int logN, const ntt_modulus_config* moduli,
                    size_t limbs, int domain, const ntt_transform_config* cfg,
                    ntt_plan** out);
int ntt_plan_destroy(ntt_plan* plan);
int ntt_plan_get_config(const ntt_plan* plan, ntt_transform_config* out);
/* Selects the kernel family (ntt_engine); NTT_ERR_CONFIG if unsupported for logN. */
int ntt_plan_set_engine(ntt_plan* plan, int engine);
/* Fused per-limb constant multiplies (the elementwise prologue/epilogue of
 * Cheddar's NTT handlers, NTT_cheddar.cu:128-229 and MultConstLazy at
 * 286-293): after this call the plan's forward transform computes
 * NTT(prologue[l] * x) and its inverse computes epilogue[l] * INTT(x), limb by
 * limb, at no extra pass (the constants fold into the R^2 / N^-1 multiplies
 * the kernels already do on load and store).  Constants are host arrays of
 * `limbs` words, reduced mod q; NULL means 1.  Synchronous; do not call while
 * launches of this plan are in flight. */
int ntt_plan_set_constants(ntt_plan* plan, const uint32_t* prologue, const uint32_t* epilogue);
/* Workgroup shape of the fused engine for both directions, per CU: 7 = 512
 * threads x 128 words each, 6 = 1024 threads x 64 words, 0 = the tuned
 * choice (default 6).  logN 17 always runs width 6 with table twiddles and
 * has no fused mod-down form (ntt_plan_forward_mod_down runs two-pass there). */
int ntt_plan_set_fused_shape(ntt_plan* plan, int eb);
/* Words per fused workgroup, log2: 16 (one workgroup per CU, every fused
 * size), 15 (two independent workgroups per CU, logN <= 15; larger sizes
 * keep 16), 0 = default (16). */
int ntt_plan_set_fused_words(ntt_plan* plan, int log2_words);
/* On-the-fly twiddles (ntt_twiddle_mode) for the plan's fused-engine launches:
 * the compressed-twiddle scheme of Cheddar (msb table of every LsbSize-th
 * twiddle, NTT_cheddar.cu:1283-1289; MultiRadixNTT_OT / MultiRadixINTT_OT in
 * the last stage group, 79-93 and 421-435).  A fused launch in this mode
 * reads 2^10 + N/64 table words per limb in place of N, at one Montgomery
 * multiply per twiddle of that round; the shape is 1024 threads x 64 words,
 * with the plan's arithmetic (lazy when every q < 2^30).  Two-pass launches
 * and the mod-down form read the table as before.  NTT_TWIDDLE_AUTO (the
 * default) is the tuned choice per direction, untuned on the fly for the
 * inverse and table twiddles for the forward. */
int ntt_plan_set_twiddle_mode(ntt_plan* plan, int mode);
/* Arithmetic of the fused kernels, both directions: 1 = lazy reduction (used
 * only when every q < 2^30), 0 = canonical, -1 = the tuned choice (default
 * lazy).  Outputs are canonical either way. */
int ntt_plan_set_fused_arith(ntt_plan* plan, int lazy);
/* The fused kernel variant (shape, lazy, words per workgroup as log2,
 * on-the-fly twiddles) a run of `count` operands uses in `direction`; any
 * output may be NULL. */
int ntt_plan_get_fused_variant(const ntt_plan* plan, size_t count, int direction, int* eb, int* lazy,
                               int* log2_words, int* on_the_fly);
/* Engine a launch of `count` operands would use (AUTO resolved). */
int ntt_plan_get_engine(const ntt_plan* plan, size_t count);
/* Engine a launch of `count` operands would use in `direction` (the
 * autotuner picks each direction's engine; ntt_plan_get_engine reports the
 * forward's). */
int ntt_plan_get_engine_for(const ntt_plan* plan, size_t count, int direction);
int ntt_plan_forward(ntt_plan* plan, uint32_t* A_dev, void* stream);
int ntt_plan_inverse(ntt_plan* plan, uint32_t* A_dev, void* stream);
/* Same transform over `count` operands laid out back to back
 * (operand k at A_dev + k * limbs * N) in one launch per phase. */
int ntt_plan_forward_many(ntt_plan* plan, uint32_t* A_dev, size_t count, void* stream);
int ntt_plan_inverse_many(ntt_plan* plan, uint32_t* A_dev, size_t count, void* stream);
/* Out-of-place forms: read `count` operands at src_dev, write their
 * transforms to dst_dev (same layout); src_dev is left unchanged unless it
 * equals dst_dev (then the call is the in-place one).  The two ranges must be
 * the same or disjoint.  The src -> dst form of Cheddar's NTTForModUp
 * (NTT_cheddar.cu:815-888: NTTPhase1 reads `src`, writes `dst`) and
 * INTTForModDown (974-1128: dst = INTT(src) * const, with the plan's epilogue
 * constants of ntt_plan_set_constants).  A limb range that ModUp skips is a
 * plan over the remaining moduli, called on the matching limb offsets. */
int ntt_plan_forward_to(ntt_plan* plan, uint32_t* dst_dev, const uint32_t* src_dev, size_t count,
                        void* stream);
int ntt_plan_inverse_to(ntt_plan* plan, uint32_t* dst_dev, const uint32_t* src_dev, size_t count,
                        void* stream);
/* Per-limb factors f_l (host array of `limbs` words, reduced mod q_l) of
 * ntt_plan_forward_mod_down, e.g. P^-1 mod q_l; NULL releases them.
 * Synchronous; the first call allocates `limbs` words of device memory. */
int ntt_plan_set_mod_down_factors(ntt_plan* plan, const uint32_t* factors);
/* The sub operand of the mod-down form restricted to limbs
 * [sub_start, sub_end) of the plan, as Cheddar's src2 limb range
 * (src/NTT_cheddar.cu:566-590): `sub` then holds sub_end - sub_start limbs
 * per operand, limb l of the plan reads sub limb l - sub_start, and limbs
 * outside the range compute dst = (0 - NTT) * f.  sub_factors (host,
 * sub_end - sub_start values, or NULL) multiplies each sub limb first, as
 * Cheddar's src2_padding (579-584).  (0, limbs, NULL) is the default: every
 * limb, no multiplier. */
int ntt_plan_set_mod_down_sub(ntt_plan* plan, uint32_t sub_start, uint32_t sub_end, const uint32_t* sub_factors);
/* NTTForModUp with a skipped limb range (src/NTT_cheddar.cu:815-887): the
 * forward transform of every limb of each of `count` operands except limbs
 * [skip_start, skip_end), in one launch; dst limbs in that range are left
 * untouched.  src NULL or equal to dst: in place; otherwise src has dst's
 * layout. */
int ntt_plan_forward_skip(ntt_plan* plan, uint32_t* dst_dev, const uint32_t* src_dev, uint32_t skip_start,
                          uint32_t skip_end, size_t count, void* stream);
/* INTTForModDown over a limb subset of its source (src/NTT_cheddar.cu:974-1049,
 * the np_non_intt case): dst (count operands of the plan's limbs) =
 * INTT of limbs [src_first, src_first + limbs) of `count` source operands of
 * src_limbs limbs each, times the plan's epilogue constants; src must not
 * overlap dst. */
int ntt_plan_inverse_from(ntt_plan* plan, uint32_t* dst_dev, const uint32_t* src_dev, uint32_t src_limbs,
                          uint32_t src_first, size_t count, void* stream);
/* dst = (sub - NTT(src)) * f_l, limb by limb, in the forward transform's
 * output order and domain (sub holds canonical residues in that order and
 * domain): the NTT and the subtract-and-scale of Cheddar's NTTForModDown
 * (NTT_cheddar.cu:889-972; the epilogue of NTTPhase2ForModDown, 573-599) in
 * one pass over the data.  src may equal dst (in place); sub must be disjoint
 * from dst (the reference asserts dst != src2, NTT_cheddar.cu:910-911).
 * NTT_ERR_INVALID_ARGUMENT until ntt_plan_set_mod_down_factors was called. */
int ntt_plan_forward_mod_down(ntt_plan* plan, uint32_t* dst_dev, const uint32_t* src_dev,
                              const uint32_t* sub_dev, size_t count, void* stream);
/* One of the two passes alone (pass 0 runs first: the column pass of a
 * forward transform, the row pass of an inverse one); the two in order equal
 * the whole transform.  For per-kernel timing, and for callers that put work
 * between the passes (e.g. a transpose across devices).  An empty pass
 * (radix_stages 0) is a no-op. */
int ntt_plan_run_pass(ntt_plan* plan, uint32_t* A_dev, size_t count, int direction, int pass,
                      void* stream);

/* ---------------------------------------------------------------------------
 * 64-bit words: the word width of Cheddar's NTTHandler<uint64_t>
 * (NTT_cheddar.cu:1313) for primes beyond 2^31; the reference's own kernels
 * are 32-bit.  Primes q < 2^62 with q = 1 mod 2N; `root` is any quadratic
 * non-residue mod q (a primitive root is one), psi = root^((q-1)/2N).
 * Layout A[limb * N + i] of uint64 words, in place.  Normal domain in and
 * out: forward maps natural-order coefficients to the bit-reversed
 * evaluations at psi^(2 brv(k) + 1); inverse maps back, N^-1 included.
 * 2 <= logN <= 17.  Two passes (column, row) of LDS-staged radix-2 stages.
 * ------------------------------------------------------------------------- */
typedef struct {
    uint64_t q;
    uint64_t root;
} ntt_modulus_config64;
typedef struct ntt_plan64 ntt_plan64;
/* The `count` largest primes q < min(bound, 2^62) with q = 1 mod 2^(logN+1),
 * descending, each with its smallest quadratic non-residue.  Returns the
 * number written. */
int ntt_generate_moduli64(int logN, uint64_t bound, int count, ntt_modulus_config64* out);
/* Builds and uploads the plan's Montgomery-form (R = 2^64) tables; checks
 * every modulus like ntt_plan_create. */
int ntt_plan64_create(ntt_context* ctx, int logN, const ntt_modulus_config64* moduli, size_t limbs,
                      ntt_plan64** out);
int ntt_plan64_destroy(ntt_plan64* plan);
/* `count` operands back to back (operand k at A_dev + k * limbs * N). */
int ntt_plan64_forward(ntt_plan64* plan, uint64_t* A_dev, size_t count, void* stream);
int ntt_plan64_inverse(ntt_plan64* plan, uint64_t* A_dev, size_t count, void* stream);

/* ---------------------------------------------------------------------------
 * Autotuner.  No counterpart in the reference, which fixes its phase split at
 * compile time (DefaultNTTLaunchConfig<LogN>, src/ntt_config.h:62-73).
 * Measures every valid (radix split, stage_merging, warp_batching) per pass
 * and direction, then the operands-per-chunk of multi-operand launches, for
 * `limbs` x `count` operands of size 2^logN; stores the winner in the context
 * for launches of that size (plans created without an explicit config use
 * it).  Allocates a scratch operand of count * limbs * N words.
 * ------------------------------------------------------------------------- */
typedef struct {
    ntt_transform_config cfg;
    size_t chunk;               /* operands per chunk, 0 = whole launch */
    double fwd_ms, inv_ms;      /* tuned forward / inverse time of the launch */
    double default_fwd_ms, default_inv_ms;   /* same with ntt_default_transform_config */
    int candidates;             /* configurations timed */
    int engine;                 /* ntt_engine chosen (TWO_PASS or FUSED) */
    double fused_fwd_ms, fused_inv_ms;       /* best fused engine times (0 if not applicable) */
    int fused_stagger;          /* start stagger of the fused engine (not searched: 0) */
    int fused_eb;               /* fused forward shape: 7 = 512 threads x 128 words, 6 = 1024 x 64 */
    int fused_eb_inv;           /* fused inverse shape */
    int fused_lazy;             /* fused forward arithmetic: 1 lazy, 0 canonical */
    int fused_lazy_inv;         /* fused inverse arithmetic: 1 lazy, 0 canonical */
    int fused_words_log2;       /* fused forward words per workgroup (log2): 16 or 15 */
    int fused_words_log2_inv;   /* fused inverse words per workgroup (log2) */
    int fused_ot;               /* fused forward twiddles: 1 on the fly, 0 table */
    int fused_ot_inv;           /* fused inverse twiddles: 1 on the fly, 0 table */
    int engine_inv;             /* ntt_engine chosen for the inverse (engine: the forward's) */
} ntt_tune_result;

int ntt_autotune(ntt_context* ctx, int logN, const ntt_modulus_config* moduli, size_t limbs,
                 size_t count, int domain, int reps, ntt_tune_result* out);
int ntt_context_set_tuning(ntt_context* ctx, int logN, size_t total_limbs,
                           const ntt_transform_config* cfg, size_t chunk, int engine);
int ntt_context_clear_tuning(ntt_context* ctx);
int ntt_context_save_tuning(const ntt_context* ctx, const char* path);
int ntt_context_load_tuning(ntt_context* ctx, const char* path);

/* ---------------------------------------------------------------------------
 * Batches: many independent transforms of uneven size, limb count, modulus
 * set and direction, launched together.  No counterpart in the reference,
 * whose main() runs one fixed-size transform at a time
 * (src/v04_montgomery.cu:969-1062).  Jobs of one (logN, direction) with
 * 10 <= logN <= 17 run as ONE persistent fused kernel that walks a
 * device-resident job table; jobs of other sizes run, per (logN, direction,
 * domain), as ONE two-pass transform over all their limbs through
 * device-resident per-limb operand pointers.
 * Groups are spread over forked streams so they overlap on the device.  The
 * whole launch sequence can be captured once into a hipGraph and replayed.
 * Twiddle tables come from the context's cache (shared across jobs).
 * ------------------------------------------------------------------------- */
typedef struct ntt_batch ntt_batch;

typedef struct {
    uint32_t* data;                       /* device operand, limbs x 2^logN words, in place */
    const ntt_modulus_config* moduli;     /* host array of `limbs` moduli (read at create) */
    uint32_t limbs;
    int32_t logN;
    int32_t direction;                    /* ntt_direction */
    int32_t domain;                       /* ntt_domain */
} ntt_batch_job;

/* Validates every job and builds (or reuses) per-modulus-set constants. */
int ntt_batch_create(ntt_context* ctx, const ntt_batch_job* jobs, size_t njobs, ntt_batch** out);
int ntt_batch_destroy(ntt_batch* batch);
/* Enqueues every job on `stream` (and on the batch's own streams, joined back
 * into `stream`); asynchronous. */
int ntt_batch_run(ntt_batch* batch, void* stream);
/* Captures ntt_batch_run into a hipGraph (replacing an earlier capture). */
int ntt_batch_capture(ntt_batch* batch, void* stream);
/* Launches the captured graph on `stream` (captures first if needed). */
int ntt_batch_replay(ntt_batch* batch, void* stream);
/* Points job `index` at another operand of the same shape.  Every job reads
 * its operand address from device memory (the fused job table or the per-limb
 * pointer array), so a captured graph stays valid.  Synchronous with respect
 * to the device. */
int ntt_batch_set_data(ntt_batch* batch, size_t index, uint32_t* data);
/* fused_groups: persistent fused launches; two_pass_jobs: jobs on two-pass
 * kernels; streams: concurrent streams used. */
int ntt_batch_info(const ntt_batch* batch, int* fused_groups, int* two_pass_jobs, int* streams);

/* ---------------------------------------------------------------------------
 * Allocation-free batch launch.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It runs the same
 * kernels as ntt_batch_run (one persistent fused kernel per (logN,
 * direction) with 10 <= logN <= 17, one two-pass transform per other
 * (logN, direction)), Montgomery variant, over jobs that bring the device
 * tables of the reference's launchers (src/v04_montgomery.cu:603-658), and
 * calls no hipMalloc or hipFree: the job tables those kernels read live in a
 * fixed device array of the library (about 0.75 MiB per GPU, held until the
 * process ends), sized for at most NTT_STATIC_MAX_JOBS jobs and
 * NTT_STATIC_MAX_LIMBS limbs in total per call (NTT_ERR_SCRATCH_LIMIT
 * beyond).  Every call on one GPU shares that array: calls that overlap on a
 * device (other streams or threads) race and must be serialised by the
 * caller.  Asynchronous on `stream`; canonical (not lazy) arithmetic.
 * ------------------------------------------------------------------------- */
#define NTT_STATIC_MAX_JOBS 1024
#define NTT_STATIC_MAX_LIMBS 16384

typedef struct {
    uint32_t* data;               /* device operand, limbs x 2^logN words, in place, 16-byte aligned */
    const uint32_t* twiddles;     /* device, limbs x 2^logN Montgomery-form table, limb-major:
                                     psi^brv(i) forward, psi^-brv(i) inverse (v04:909-915) */
    const uint32_t* moduli;       /* device, limbs: q */
    const uint32_t* q_inv;        /* device, limbs: -q^-1 mod 2^32 */
    const uint32_t* scale;        /* device, limbs: R^2 mod q forward, to_mont(N^-1) inverse */
    uint32_t limbs;
    int32_t logN;
    int32_t direction;            /* ntt_direction */
} ntt_table_job;

int ntt_batch_run_static(const ntt_table_job* jobs, size_t njobs, void* stream);

/* ---------------------------------------------------------------------------
 * Serving: a request loop with continuous batching.  No counterpart in the
 * reference (its main() runs one synchronous transform at a time,
 * src/v04_montgomery.cu:969-1062).  A scheduler thread owns an operand arena
 * of hbm_budget_bytes in HBM.  Submitted requests queue in arrival order; a
 * request is admitted when its operand fits in the arena (first fit,
 * strictly FIFO, so nothing starves), its operand is placed there (copied
 * from the given source or generated from a seed), and admitted requests are
 * launched together on the first free launch slot (one stream each): per
 * (logN, direction) group one persistent fused kernel, two-pass kernels for
 * other sizes.  Requests that arrive while launches are in flight are
 * admitted as arena space and slots free up and run concurrently with them.
 * Results stay resident in the arena until fetched or released.  Twiddle
 * tables stay resident in the context's cache.
 * ------------------------------------------------------------------------- */
typedef struct ntt_server ntt_server;

typedef struct {
    size_t hbm_budget_bytes;        /* operand arena size */
    int slots;                      /* launches in flight at once, 1..8 */
    int max_jobs_per_launch;        /* 1..4096 */
} ntt_server_config;

typedef struct {
    int32_t logN;
    uint32_t limbs;
    int32_t direction;              /* ntt_direction */
    int32_t domain;                 /* ntt_domain */
    /* host array of `limbs` moduli, or NULL: the `limbs` largest primes
     * q < 2^30 with q = 1 mod 2N (ntt_generate_moduli) */
    const ntt_modulus_config* moduli;
    /* operand source on the host or on the device (one of them, or neither);
     * it must stay valid until the request is done */
    const uint32_t* src_host;
    const uint32_t* src_dev;
    uint64_t seed;                  /* both sources NULL: uniform residues (ntt_rt_fill_uniform) */
    int32_t auto_release;           /* 1: free the operand when done (result not kept); the
                                       request stays known to state/wait until released */
} ntt_request;

typedef enum {
    NTT_REQ_QUEUED = 0, NTT_REQ_ADMITTED = 1, NTT_REQ_RUNNING = 2, NTT_REQ_DONE = 3,
    NTT_REQ_FAILED = 4, NTT_REQ_UNKNOWN = 5
} ntt_request_state;

typedef struct {
    uint64_t submitted, completed, launches, fused_launches, two_pass_launches;
    double mean_jobs_per_launch;
    size_t arena_bytes, arena_peak_bytes;
    int max_slots_busy;             /* most launches seen in flight at once */
} ntt_server_stats;

int ntt_server_create(ntt_context* ctx, const ntt_server_config* cfg, ntt_server** out);
/* Stops the scheduler after the work in flight has finished; requests not
 * yet admitted fail with NTT_ERR_BUSY. */
int ntt_server_destroy(ntt_server* server);
/* Queues a request (arrival time = now); NTT_ERR_BUDGET if its operand can
 * never fit the arena. */
int ntt_server_submit(ntt_server* server, const ntt_request* req, uint64_t* id);
int ntt_server_state(ntt_server* server, uint64_t id, int* state);
/* Blocks until request `id` is done (or failed: its status is returned). */
int ntt_server_wait(ntt_server* server, uint64_t id);
/* Copies a finished request's result (limbs x 2^logN words) to `dst` (device
 * memory if dst_is_device, else host) and, if release, frees its operand.
 * The result stays allocated while any fetch of it is copying: a release (or
 * a fetch with release) that arrives meanwhile takes effect when the last
 * such copy ends.  After a release, fetch returns NTT_ERR_NOT_FOUND. */
int ntt_server_fetch(ntt_server* server, uint64_t id, uint32_t* dst, int dst_is_device, int release);
int ntt_server_release(ntt_server* server, uint64_t id);
int ntt_server_stats_get(ntt_server* server, ntt_server_stats* out);
/* Builds the plans and twiddle tables for requests of (logN, limbs) with
 * the default moduli ahead of time, so first requests of a shape do no
 * host-side setup. */
int ntt_server_prepare(ntt_server* server, int logN, uint32_t limbs, int domain);
/* Trace driver: submits request i (default moduli, seeded operand,
 * auto-release) at arrival_s[i] seconds after the start, waits for all, and
 * writes each latency (completion seen by the scheduler - arrival) to
 * latency_s[i]; *wall_s = first arrival to last completion. */
int ntt_server_run_trace(ntt_server* server, size_t n, const double* arrival_s, const int32_t* logN,
                         const uint32_t* limbs, const int32_t* direction, const uint64_t* seed,
                         int domain, double* latency_s, double* wall_s);

#ifdef __cplusplus
}
#endif

#endif /* NTT_AMD_H */
