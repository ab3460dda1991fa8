/*
 * ldpc_mi355x.h -- C ABI of the MI355X-native 8-bit fixed-point Min-Sum LDPC
 * decoder (libldpc_mi355x.so, built for gfx950).
 *
 * This is the drop-in boundary for the reference's GPU decoding path.  Each
 * entry point names the reference interface it replaces (paths relative to the
 * reference repository root).  Plain C types only; HIP streams are passed as
 * `void *` (a hipStream_t), device pointers as plain pointers.
 *
 * Conventions kept from the reference:
 *   - `nb_frames` counts 32-bit WORDS; one word packs 4 frames (frames = 4*W)
 *     (src/main.cpp:210,214,309; src/decoder_template/CGPUDecoder.cpp:55-57).
 *   - an LLR/posterior/message is an int8; a positive value means bit 1, so
 *     BPSK maps bit 0 to -1 (src/awgn_channel/CChanel_AWGN_SIMD.cu:21-22) and
 *     the hard decision is (post > 0) (src/transpose/GPU_Transpose_uint8.cu:29).
 *   - host input is frame-major int8 in[f*N + v], f in [0, 4W); the OMS_v2
 *     ("xMS") variant consumes it as the device layout post[v*W + w], byte k =
 *     frame 4w+k (src/decoder_oms_v2/CGPU_Decoder_MS_SIMD_v2.cu:187).
 *   - output is frame-major bytes out[f*N + v] in {0, 1}.
 * Differences (documented, deliberate): errors are returned as status codes
 * (the reference prints and calls exit(0), src/custom_api/custom_cuda.cu:5-17);
 * iterations < 1 is rejected (the reference wraps to ~2^32 iterations,
 * src/decoder_oms/cuda/CUDA_OMS_SIMD.cu:132,138); W need not be a multiple of
 * 128 (src/decoder_oms/CGPU_Decoder_OMS_SIMD.cu:101-105), only of the block.
 */
#ifndef LDPC_MI355X_H
#define LDPC_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */
enum {
    LDPC_OK = 0,
    LDPC_ERR_ARG = -1,      /* invalid argument / precondition violated */
    LDPC_ERR_HIP = -2,      /* HIP runtime error (message via ldpc_last_error) */
    LDPC_ERR_NOMEM = -3,    /* allocation failed */
    LDPC_ERR_NODEV = -4,    /* no usable gfx950 device */
    LDPC_ERR_STATE = -5     /* call order (e.g. iterate before load) */
};
/* Text of the last error on this thread (replaces ERROR_CHECK's printf,
 * src/custom_api/custom_cuda.cu:5-17). */
const char *ldpc_last_error(void);
/* Library version string. */
const char *ldpc_version(void);

/* ---- variants (one per reference decoder object) ------------------------ */
enum {
    LDPC_MS = 0,     /* CGPU_Decoder_MS_SIMD    src/decoder_ms/CGPU_Decoder_MS_SIMD.cu      */
    LDPC_OMS = 1,    /* CGPU_Decoder_OMS_SIMD   src/decoder_oms/CGPU_Decoder_OMS_SIMD.cu    */
    LDPC_OMS_V2 = 2, /* CGPU_Decoder_MS_SIMD_v2 src/decoder_oms_v2/CGPU_Decoder_MS_SIMD_v2.cu */
    LDPC_NMS = 3,    /* CGPU_Decoder_NMS_SIMD   src/decoder_nms/CGPU_Decoder_NMS_SIMD.cu    */
    LDPC_2NMS = 4,   /* CGPU_Decoder_2NMS_SIMD  src/decoder_2nms/CGPU_Decoder_2NMS_SIMD.cu  */
    LDPC_LOMS = 5,   /* learned-offset OMS (int8 table [iteration][check degree])   */
    LDPC_LNMS = 6    /* learned-normalised MS (table value a: min(floor(m*a/8), 31)) */
};

/* ---- parity-check codes (replaces the compile-time tables of src/matrix:
 * PosNoeudsVariable[_M] + DEG_1/DEG_2/DEG_*_COMPUTATIONS macros,
 * src/matrix/4000x2000/constantes_gpu.h:6-38, constantes_decoder.h:3) ----- */
typedef struct ldpc_code ldpc_code;

/* Table form: check c (table order) has degree deg[c] and its variables are
 * pos[off_c .. off_c+deg[c]) with off_c = sum of earlier degrees.  The
 * serial layered decoder processes checks in exactly this order.
 * q1_split: first table index of the reference's second degree group (OMS
 * first-iteration no-clamp quirk, src/decoder_oms/cuda/CUDA_OMS_SIMD.cu:115-116);
 * pass -1 to use the end of the leading run of equal degrees. */
ldpc_code *ldpc_code_create(int n_vars, int n_checks, const int *deg,
                            const uint32_t *pos, int q1_split);
/* The reference's macro form: deg1 x count1 checks then deg2 x count2
 * (src/matrix/1944x972/constantes_gpu.h:6-20). count2 may be 0. */
ldpc_code *ldpc_code_create_grouped(int n_vars, int deg1, int count1, int deg2,
                                    int count2, const uint32_t *pos);
/* Parse a reference table file ("/" "* msg = i, deg = d *" "/  v0, v1, ..."),
 * e.g. src/matrix/4000x2000/constantes_decoder.h, read from `path` at run time. */
ldpc_code *ldpc_code_load_table(const char *path, int n_vars);
/* Seeded generators (deterministic across platforms).
 * Regular (dv, dc): random socket permutation, 4-cycles and repeated
 * variables rejected (the (3,6) N = 4000 code: n_vars 4000, dv 3, dc 6). */
ldpc_code *ldpc_code_regular(int n_vars, int dv, int dc, uint64_t seed);
/* Quasi-cyclic: mb x nb base graph, circulant size z, dual-diagonal parity
 * part on the last mb columns, info entries present with probability p (an
 * info column is redrawn until it has at least two entries), shifts uniform
 * in [0, z-1]. */
ldpc_code *ldpc_code_qc(int mb, int nb, int z, double p, uint64_t seed);
/* IRA staircase code: k_info info bits, n_hi of them of degree d_hi, the rest
 * of degree d_lo, k_info checks, parity degree 2 (staircase), info edges
 * spread evenly over checks by a seeded permutation. */
ldpc_code *ldpc_code_ira(int k_info, int n_hi, int d_hi, int d_lo, uint64_t seed);
/* New code with the same H whose table order lists checks colour class by
 * colour class, so the serial schedule has few, evenly filled dependency
 * levels.  Colouring of the check conflict graph (checks sharing a variable
 * conflict): DSatur order, each check in the least-filled colour its
 * neighbours leave free, starting from as many colours as the largest
 * variable degree; then iterated greedy and tabu search cut the colour count
 * (the tabu stage for conflict tables up to 2^22 checks x colours), and
 * checks move from over-full classes into free smaller ones; deterministic. */
ldpc_code *ldpc_code_reorder_by_colour(const ldpc_code *code);
void ldpc_code_destroy(ldpc_code *code);

/* n_vars, n_checks, n_edges, max check degree, dependency levels of the
 * serial order, q1_split. Any pointer may be NULL. */
int ldpc_code_info(const ldpc_code *code, int *n_vars, int *n_checks, int *n_edges,
                   int *max_deg, int *n_levels, int *q1_split);
/* Copy out the table (deg[n_checks], pos[n_edges]). */
int ldpc_code_export(const ldpc_code *code, int *deg, uint32_t *pos);
/* The level of every check in the decoders' parallel schedule of the table
 * (level_of[n_checks], may be NULL); returns the number of levels, or a
 * negative status.  granule <= 1: dependency levels (a check's level is one
 * above the highest level of the earlier checks sharing a variable with it).
 * granule > 1: the same number of levels, checks with slack moved so that
 * level sizes are multiples of granule where the dependencies allow; the word
 * decoders use 64 / words_per_block (full 64-lane waves).  Checks sharing a
 * variable keep their table order either way, so decoding is unchanged.  No
 * reference counterpart: the reference runs the table serially per word
 * (src/decoder_oms/CUDA_OMS_SIMD.cu:148-191). */
int ldpc_code_levels(const ldpc_code *code, int granule, int *level_of);

/* ---- decoder objects (replaces class CGPUDecoder,
 * src/decoder_template/CGPUDecoder.h:19-35, and its five subclasses) ------ */
typedef struct ldpc_decoder ldpc_decoder;

typedef struct {
    int words_per_block;  /* P: 0 = auto (power of two <= 16 dividing W)     */
    int threads;          /* workgroup size: 0 = auto                         */
    int early_term;       /* -1 = variant default (2NMS: on), 0 off, 1 on      */
    int post_in_lds;      /* -1 = auto (posterior block fits LDS), 0, 1        */
    int uniform;          /* single-degree kernel: -1 = auto, 0 = off, 1 = on  */
    int batch;            /* checks per lane whose loads are issued together:
                             0 = auto, 1, 2 or 4 (single-degree kernel)        */
} ldpc_tuning;

/* CGPUDecoder(nb_frames, n, k, m) + derived ctor: device buffers and the H
 * table upload (src/decoder_template/CGPUDecoder.cpp:51-75).  The code is
 * copied; `tuning` may be NULL. */
ldpc_decoder *ldpc_decoder_create(const ldpc_code *code, size_t nb_frames_words,
                                  int variant, const ldpc_tuning *tuning);
/* CGPUDecoder::initialize() (src/decoder_oms/CGPU_Decoder_OMS_SIMD.cu:92-94). */
int ldpc_decoder_initialize(ldpc_decoder *dec);
/* ~CGPUDecoder() (src/decoder_template/CGPUDecoder.cpp:78-91). */
void ldpc_decoder_destroy(ldpc_decoder *dec);

/* decode(var_nodes, Rprime_fix, nombre_iterations): synchronous host-buffer
 * decode, H2D -> interleave -> iterations -> hard decision -> D2H
 * (src/decoder_oms/CGPU_Decoder_OMS_SIMD.cu:97-149).
 * in: 4W*N int8 (frame-major; var-major words for OMS_v2); out: 4W*N bytes.
 * A batch of 32 MiB and 4096 word blocks or more is copied and decoded in 8
 * chunks whose copies overlap the other chunks' decodes (pinned host buffers
 * make the copies asynchronous); the result is the same. */
int ldpc_decoder_decode(ldpc_decoder *dec, const int8_t *in, uint8_t *out, int iterations);
/* Same on device buffers, enqueued on `stream` (NULL = the decoder's own).
 * Asynchronous: returns after the launch.  Frames whose 4-frame word block
 * does not fit in LDS (and were left to the automatic choice) are decoded,
 * when the iterations are fixed, two per 16-bit lane by the frame-queue
 * kernel with no stop test (same decisions, posterior in LDS); OMS on codes
 * with two degree groups and the per-word early stop use the global-memory
 * posterior.  Calls on one decoder must not overlap in that case. */
int ldpc_decoder_decode_device(ldpc_decoder *dec, const int8_t *d_in, uint8_t *d_out,
                               int iterations, void *stream);

/* The decoder's own HIP stream (hipStream_t) used when no stream is given. */
void *ldpc_decoder_stream(ldpc_decoder *dec);
/* Words per workgroup block, workgroup size, posterior-in-LDS flag,
 * dependency levels, single-degree kernel degree (0 = generic kernel) and
 * load batch chosen at create time. Any pointer may be NULL. */
int ldpc_decoder_config(const ldpc_decoder *dec, int *words_per_block, int *threads,
                        int *post_in_lds, int *n_levels, int *uniform_deg, int *batch);

/* Learned table for LDPC_LOMS / LDPC_LNMS: int8 [n_iters][n_degs], row =
 * absolute iteration (last row repeats), column = check degree (last column
 * repeats).  LOMS: offset in LSB, 0..127. LNMS: numerator over 8, 0..8,
 * magnitude capped at 31.  At most 64 rows and 33 columns (the frame-queue
 * decoder's limits; a rejected table leaves the decoder unchanged). */
int ldpc_decoder_set_table(ldpc_decoder *dec, const int8_t *table, int n_iters, int n_degs);

/* ---- state API (per-iteration parity; the reference keeps this state in
 * device_V / d_MSG_C_2_V, src/decoder_template/CGPUDecoder.h:21-23) ------- */
/* Load host input (same layout as decode) and reset the decode. */
int ldpc_decoder_load(ldpc_decoder *dec, const int8_t *in);
/* Run `iterations` more iterations from the current state. */
int ldpc_decoder_iterate(ldpc_decoder *dec, int iterations);
/* Posterior in the reference layout post[v*W + w] (N*W words) and messages
 * in the reference layout msg[e*W + w], e in table order (E*W words).
 * Either pointer may be NULL. */
int ldpc_decoder_get_state(ldpc_decoder *dec, uint32_t *post, uint32_t *msg);
/* Hard decision of the current posterior, frame-major bytes (4W*N). */
int ldpc_decoder_get_hard(ldpc_decoder *dec, uint8_t *out);
/* Per-word stop flags (W bytes) and iterations run since load (W ints). */
int ldpc_decoder_get_stop_info(ldpc_decoder *dec, uint8_t *stopped, int *iters_run);

/* ---- device interleave (src/transpose/GPU_Transpose_uint8.cu:80-130) and
 * inverse with hard decision (:9-78) on device buffers ------------------- */
int ldpc_interleave_device(const int8_t *d_in, uint32_t *d_words, int n_vars,
                           size_t nb_frames_words, void *stream);
int ldpc_deinterleave_hd_device(const uint32_t *d_words, uint8_t *d_out, int n_vars,
                                size_t nb_frames_words, void *stream);

/* ---- device AWGN channel (replaces CChanel_AWGN_SIMD::configure/generate,
 * src/awgn_channel/CChanel_AWGN_SIMD.cu:63-112: cuRAND + Box-Muller + a D2H copy
 * per chunk).  All-zero codeword, BPSK bit 0 -> -1, y = -1 + sigma*z; sample
 * s = frame*N + var is a pure function of (seed, snr_idx, s): Philox4x32-10
 * counter (s_lo, s_hi, snr_idx, 0x4C445043), key (seed_lo, seed_hi), two
 * 53-bit uniforms, Box-Muller cosine branch, in double precision.
 * quant 0: sat127(round(16*y/sigma^2))   (8 * the LLR 2y/sigma^2)
 * quant 1: (int8)trunc(clamp(8*y, +-31)) (CChanel_AWGN_SIMD.cu:23-24)
 * layout 0: frame-major int8 [frames][N] (the decode input);
 * layout 1: device words [N][frames/4] (OMS_v2 input / interleaved state). */
int ldpc_channel_generate(void *d_out, int layout, int n_vars, size_t frames, uint64_t first_frame,
                          uint64_t seed, uint32_t snr_idx, double sigma, int quant, void *stream);
/* sigma of CChanel_AWGN_SIMD::configure (:63-73); es_n0 != 0 reads `snr_db`
 * as Es/N0. */
double ldpc_channel_sigma(double snr_db, double rate, int es_n0);

/* ---- device error counting (replaces CErrorAnalyzer::generate,
 * src/ber_analyzer/CErrorAnalyzer.cpp:141-157, a host scan) -----------------
 * d_hard: frame-major decisions [frames][N] (all-zero codeword, so a 1 is an
 * error).  Adds to d_counters[0..2] (unsigned 64-bit, device memory): bit
 * errors in positions < n_info, frames with a bit error, frames whose
 * decisions violate a parity check of `dec`'s code.  accumulate = 0 zeroes
 * the counters first.  Per-frame outputs may be NULL. Asynchronous. */
int ldpc_decoder_count_errors(ldpc_decoder *dec, const uint8_t *d_hard, size_t frames, int n_info,
                              uint64_t *d_counters, uint32_t *d_frame_bit_errors,
                              uint8_t *d_frame_syndrome_ok, int accumulate, void *stream);

/* Training frames: per-frame Eb/N0 = lo + (hi-lo)*u(f) dB, u from the Philox
 * block (f_lo, f_hi, snr_idx, 0x534E5252); sigma = sqrt(0.5/(rate*10^(EbN0/10)));
 * frame-major int8, otherwise as ldpc_channel_generate. */
int ldpc_channel_generate_range(void *d_out, int n_vars, size_t frames, uint64_t first_frame,
                                uint64_t seed, uint32_t snr_idx, double ebn0_lo, double ebn0_hi,
                                double rate, int quant, void *stream);

/* ---- simulation of one Eb/N0 point (the loop body of src/main.cpp:237-335:
 * channel -> decode -> error count, repeated over batches; device only) --- */
typedef struct {
    uint64_t frames;             /* frames decoded (multiple of 4W)            */
    uint64_t bit_errors;         /* in the first N - #checks positions          */
    uint64_t frame_errors;
    uint64_t syndrome_failures;  /* frames whose decisions fail a parity check */
    double avg_iterations;       /* per frame (early stop counts actual ones)  */
    double decode_ms;            /* hipEvent time of the decode kernels only   */
} ldpc_sim_stats;
/* Frames i*4W .. of the stream (seed, snr_idx), so two decoders given the same
 * arguments see identical frames. */
int ldpc_simulate(ldpc_decoder *dec, double ebn0_db, int quant, uint64_t seed, uint32_t snr_idx,
                  uint64_t n_frames, int iterations, ldpc_sim_stats *out);

/* ---- persistent frame-queue decoder ("pool"; no single counterpart in the
 * reference: it generalises the per-word early stop of
 * src/decoder_2nms/cuda/CUDA_2NMS_SIMD.cu:237-244 to each frame and keeps the
 * GPU full by refilling a stopped frame's lane from a device queue) --------
 * A frame stops after the first iteration (the first included) that passes
 * its stop test, or at max_iterations; its decisions and iteration count are
 * written and its slot takes the next queued frame.  Stop tests:
 *   LDPC_STOP_SYNDROME    the decisions (post > 0) satisfy every check;
 *   LDPC_STOP_V2C_PARITY  every check had an even parity of positive v2c
 *                         values during the iteration (the reference's rule,
 *                         applied to each frame);
 *   LDPC_STOP_NONE        none: every frame runs max_iterations (its
 *                         decisions equal the fixed-iteration decoder's).
 * Variants: MS, OMS_V2, NMS, 2NMS, LOMS, LNMS, and OMS on codes with a single
 * degree group (the OMS first-iteration quirk is not per-frame).  Results
 * per frame equal a one-frame layered decode with that stop test.
 * Posterior format (tuning->post_in_lds): 1 = words of 4 frames (N * 4 * P
 * bytes of LDS per workgroup), 2 = pairs of frames (N * 2 * P bytes; for
 * long frames such as N = 64800 whose word block does not fit), -1 = words
 * when they fit, else pairs.  words_per_block then counts words or pairs. */
enum { LDPC_STOP_SYNDROME = 0, LDPC_STOP_V2C_PARITY = 1, LDPC_STOP_NONE = 2 };
typedef struct ldpc_pool ldpc_pool;
ldpc_pool *ldpc_pool_create(const ldpc_code *code, int variant, int max_iterations, int stop_rule,
                            const ldpc_tuning *tuning);
int ldpc_pool_set_table(ldpc_pool *pool, const int8_t *table, int n_iters, int n_degs);
/* Queue order of the frames of one ldpc_pool_decode / ldpc_pool_decode_graph
 * call (no counterpart in the reference, whose words all run the same fixed
 * number of iterations).  LDPC_ORDER_INPUT: frame i is queued i-th.
 * LDPC_ORDER_LEAST_RELIABLE_FIRST: frames are queued by the sum of squares of
 * their first 4096 LLRs, largest first (noisiest first, in 1024 buckets over
 * the call's own range of that sum), so the frames that need many iterations
 * start early and a call ends on short ones.  Every frame's decisions and
 * iteration count are the same in either order.  The order needs
 * 8 bytes per frame of device scratch, allocated by the decode call when the
 * call has more frames than any before (so the first call at a new largest
 * size is not capturable by the caller; ldpc_pool_decode_graph allocates
 * before its own capture); frames above 2^32 - 1 are LDPC_ERR_HIP. */
enum { LDPC_ORDER_INPUT = 0, LDPC_ORDER_LEAST_RELIABLE_FIRST = 1 };
int ldpc_pool_set_order(ldpc_pool *pool, int order);
/* Decode `frames` frame-major LLRs at d_in into d_out (decisions) and
 * d_iters (iterations per frame, may be NULL).  workgroups = 0: as many as
 * stay resident on the GPU (capped by the work).  Asynchronous on `stream`
 * (NULL = the pool's own); the queue counter is reset with a kernel, so the
 * call may be captured in a HIP graph.  Calls on one pool must not overlap:
 * they share its queue head and message buffer (use one pool per stream).
 * frames = 0 is LDPC_ERR_ARG (as is an empty step of ldpc_mixed_decode). */
int ldpc_pool_decode(ldpc_pool *pool, const int8_t *d_in, uint64_t frames, uint8_t *d_out, int *d_iters,
                     int workgroups, void *stream);
/* The same decode step captured in a hipGraph (the reset kernel and the
 * frame-queue kernel) on first use and replayed on `stream` while the buffer
 * arguments, the frame count, the grid, the table and max_iterations stay the
 * same; a change recaptures (the per-batch decode() call of the reference's
 * simulation loop, src/main.cpp:266-268, as a replayed graph).  Results equal
 * ldpc_pool_decode's; the same no-overlap rule holds. */
int ldpc_pool_decode_graph(ldpc_pool *pool, const int8_t *d_in, uint64_t frames, uint8_t *d_out, int *d_iters,
                           int workgroups, void *stream);
/* One Eb/N0 point: device channel (frames 0 .. n_frames-1 of the stream
 * (seed, snr_idx), the same frames ldpc_simulate draws) -> pool decode ->
 * device error counting, batch_frames frames per launch (0 = 2^18).
 * avg_iterations is per frame; decode_ms covers the decode launches only. */
int ldpc_pool_simulate(ldpc_pool *pool, double ebn0_db, int quant, uint64_t seed, uint32_t snr_idx,
                       uint64_t n_frames, uint64_t batch_frames, ldpc_sim_stats *out);
/* Autotuner (no counterpart in the reference): times the frame-queue
 * decoder on sample_frames channel frames at ebn0_db (quantizer `quant`, as
 * ldpc_channel_generate; iteration counts depend on it) for every words-per-
 * workgroup P (posterior block N*P*4 bytes of LDS), workgroup size (whole
 * waves: 64, 128, 192, 256, 320, 384, 512, 768, 1024, as the kernel allows)
 * and, on the single-degree path, items in flight per thread (1, 2, 4); both
 * the 4-frame word and the 2-frame pair formats; writes the fastest
 * into *best (words_per_block, threads, uniform, batch set; other fields -1).
 * Learned variants are timed with a constant table. */
int ldpc_pool_autotune(const ldpc_code *code, int variant, int max_iterations, int stop_rule, double ebn0_db,
                       int quant, uint64_t sample_frames, ldpc_tuning *best, double *best_ms, int *n_tried);
/* Words (pairs) per workgroup, workgroup size, resident workgroups, the
 * degree of the single-degree path (0 = generic path), its items per thread,
 * and whether the posterior holds 2-frame pairs. */
int ldpc_pool_config(const ldpc_pool *pool, int *words_per_block, int *threads, int *max_workgroups,
                     int *uniform_deg, int *batch, int *pairs);
void *ldpc_pool_stream(ldpc_pool *pool);
void ldpc_pool_destroy(ldpc_pool *pool);

/* ---- mixed-code decode step (no single counterpart in the reference, whose
 * code is fixed per binary: src/matrix/code.h:1) ------------------------------
 * One frame-queue decoder per code (tunings[c], e.g. from ldpc_pool_autotune,
 * or NULL = defaults), all with the same variant, iteration limit and stop
 * test.  A step
 * decodes `frames` frames held in one byte buffer: frame f belongs to code
 * d_code_ids[f] and sits at d_in + d_offsets[f] (its N bytes of LLRs); its
 * decisions go to d_out + d_offsets[f] and its iteration count to d_iters[f]
 * (0 for a frame whose code id is not a code of this object).  The step is
 * asynchronous on `stream` (NULL = the null stream); use_graph = 1 captures it
 * into a hipGraph on first use and replays the graph while the six buffer
 * arguments stay the same (a change, or a new table, recaptures).
 * max_frames sizes the per-code queues (0 = on first use). */
typedef struct ldpc_mixed ldpc_mixed;
ldpc_mixed *ldpc_mixed_create(const ldpc_code *const *codes, int n_codes, int variant, int max_iterations,
                              int stop_rule, const ldpc_tuning *tunings, uint64_t max_frames);
int ldpc_mixed_set_table(ldpc_mixed *mixed, int code_index, const int8_t *table, int n_iters, int n_degs);
int ldpc_mixed_decode(ldpc_mixed *mixed, const int8_t *d_in, const uint64_t *d_offsets, const uint8_t *d_code_ids,
                      uint64_t frames, uint8_t *d_out, int *d_iters, int use_graph, void *stream);
int ldpc_mixed_config(const ldpc_mixed *mixed, int code_index, int *words_per_block, int *threads,
                      int *max_workgroups, int *uniform_deg, int *batch, int *pairs);
void ldpc_mixed_destroy(ldpc_mixed *mixed);

/* The mixed-code step without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, semantics and results as ldpc_mixed_decode with
 * use_graph = 0, but the step's scratch (one queue counter per code and one
 * frame list of `frames` entries per code) lives in a static device array of
 * LDPC_MIXED_NOALLOC_ENTRIES + 256 uint32 (8 MiB + 1 KiB) that each GPU holds
 * until the process ends, instead of memory the object allocates; the call
 * makes no hipMalloc/hipFree and instantiates no graph.  frames * n_codes
 * must not exceed LDPC_MIXED_NOALLOC_ENTRIES (else LDPC_ERR_ARG, nothing
 * enqueued); the current device must be the one the object was created on.
 * Every such call on one GPU shares that GPU's array: calls that overlap on
 * one GPU (other streams, other objects, other host threads) race and must be
 * ordered by the caller. */
#define LDPC_MIXED_NOALLOC_ENTRIES (1u << 21)
int ldpc_mixed_decode_noalloc(ldpc_mixed *mixed, const int8_t *d_in, const uint64_t *d_offsets,
                              const uint8_t *d_code_ids, uint64_t frames, uint8_t *d_out, int *d_iters,
                              void *stream);

/* ---- learned tables: float shadow decoder trained on the GPU ----------------
 * (no counterpart in the reference).  The shadow runs the integer decoder's
 * layered schedule in float LSB units with a per-iteration, per-check-degree
 * offset (mode OFFSET: mag = clamp(min - beta, 0, 31)) or scale (mode SCALE:
 * mag = clamp(alpha * min, 0, 31)); the loss is the mean over iterations and bits of
 * softplus(posterior / 8) on all-zero codewords.  Quantised tables drive
 * LDPC_LOMS (offsets in LSB) and LDPC_LNMS (round(8 * alpha)). */
enum { LDPC_LEARN_OFFSET = 0, LDPC_LEARN_SCALE = 1 };
typedef struct ldpc_trainer ldpc_trainer;
ldpc_trainer *ldpc_trainer_create(const ldpc_code *code, int mode, int iterations, int n_degs,
                                  size_t batch_frames);
void ldpc_trainer_destroy(ldpc_trainer *tr);
/* Loss and gradient [iterations][n_degs] for `params` on device LLRs
 * (frame-major, batch_frames frames; NULL = the trainer's own channel
 * buffer).  grad and d_post_out (posterior after the last iteration,
 * [frames][N] floats on the device) may be NULL. */
int ldpc_trainer_loss_grad(ldpc_trainer *tr, const float *params, const int8_t *d_llr, float *loss,
                           float *grad, float *d_post_out);
typedef struct {
    int steps;                 /* Adam steps, one fresh batch each              */
    double lr, beta1, beta2, eps;
    double ebn0_lo, ebn0_hi;   /* per-frame Eb/N0 range (dB)                    */
    double rate;               /* <= 0: (N - #checks) / N                       */
    uint64_t seed, first_frame;
    int quant;                 /* channel quantizer (0 or 1)                    */
    double pmin, pmax;         /* parameter box                                 */
} ldpc_learn_config;
/* Adam on `params` (in/out); loss_history[steps] may be NULL. */
int ldpc_trainer_fit(ldpc_trainer *tr, const ldpc_learn_config *cfg, float *params,
                     float *loss_history);

/* ---- device memory helpers (plumbing for callers without a HIP runtime
 * binding; replaces CUDA_MALLOC_DEVICE/CUDA_MALLOC_HOST,
 * src/custom_api/custom_cuda.h:24-40, and CTrame's pinned buffers,
 * src/trame/CTrame.cpp:38-41) ------------------------------------------ */
void *ldpc_device_alloc(size_t bytes);
void ldpc_device_free(void *p);
void *ldpc_host_alloc_pinned(size_t bytes);
void ldpc_host_free_pinned(void *p);
int ldpc_memcpy_h2d(void *dst, const void *src, size_t bytes);
int ldpc_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Device-to-device copy (a plain copy of known size: HBM counter calibration). */
int ldpc_memcpy_d2d(void *dst, const void *src, size_t bytes);
int ldpc_device_synchronize(void);
/* Number of visible gfx950 devices (0 without a GPU). */
int ldpc_device_count(void);

/* ---- timer (replaces CTimer's cudaEvent pair, src/timer/CTimer.cu:4-91) --- */
typedef struct ldpc_timer ldpc_timer;
ldpc_timer *ldpc_timer_create(void *stream);
int ldpc_timer_start(ldpc_timer *t);
int ldpc_timer_stop(ldpc_timer *t);
/* Elapsed milliseconds between start and stop (synchronises on stop). */
double ldpc_timer_elapsed_ms(ldpc_timer *t);
void ldpc_timer_destroy(ldpc_timer *t);

#ifdef __cplusplus
}
#endif
#endif
