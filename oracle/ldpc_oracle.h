/*
 * ldpc_oracle.h -- CPU restatement of the reference's 8-bit fixed-point
 * Min-Sum decoding path, used ONLY as a checker by the tests, by
 * __graft_entry__.smoke() and as bench.py's cpu_baseline leg.
 *
 * TEST INFRASTRUCTURE: the product (the HIP library under
 * fast_ldpc_decoder_for_gpu_fixed_amd/) never links, loads or calls this code.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).
 */
#ifndef LDPC_ORACLE_H
#define LDPC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Variants, numbered like the product's C-ABI (include/ldpc_mi355x.h). */
enum {
    ORC_MS = 0,      /* src/decoder_ms/cuda/CUDA_MS_SIMD.cu:25-248            */
    ORC_OMS = 1,     /* src/decoder_oms/cuda/CUDA_OMS_SIMD.cu:25-248          */
    ORC_OMS_V2 = 2,  /* src/decoder_oms_v2/cuda/CUDA_OMS_SIMD_v2.cu:30-575    */
    ORC_NMS = 3,     /* src/decoder_nms/cuda/CUDA_NMS_SIMD.cu:25-292          */
    ORC_2NMS = 4,    /* src/decoder_2nms/cuda/CUDA_2NMS_SIMD.cu:25-292        */
    ORC_LOMS = 5,    /* learned-offset OMS: table[iteration][check degree]    */
    ORC_LNMS = 6     /* learned-normalised MS: min(m * table / 8, 31)         */
};

/* ---- per-byte SIMD ops (semantics: src/decoder_ms/cuda/simd_functions.h:89-131) */
uint32_t orc_vaddss4(uint32_t a, uint32_t b);
uint32_t orc_vsubss4(uint32_t a, uint32_t b);
uint32_t orc_vsubus4(uint32_t a, uint32_t b);
uint32_t orc_vminu4(uint32_t a, uint32_t b);
uint32_t orc_vmaxu4(uint32_t a, uint32_t b);
uint32_t orc_vabs4(uint32_t a);
uint32_t orc_vneg4(uint32_t a);
uint32_t orc_vcmpgts4(uint32_t a, uint32_t b);
uint32_t orc_vcmpeq4(uint32_t a, uint32_t b);
uint32_t orc_vcmpne4(uint32_t a, uint32_t b);
uint32_t orc_vsetgts4(uint32_t a, uint32_t b);

/* ---- serial layered decode over words (4 frames per uint32) -------------
 * deg[c]   : degree of check c, table order            (C checks)
 * pos[e]   : variable of edge e, table order            (E = sum deg)
 * post     : posterior, var-major [N][W] words          (in/out)
 * msg      : check->variable messages, table-order edge-major [E][W] (in/out)
 * iters    : iterations to run in this call (>= 1)
 * first    : 1 if this call starts a decode (first iteration reads no message)
 * q1_split : table index where the reference's second degree group starts
 *            (OMS first iteration skips the 31 clamp there, quirk of
 *            src/decoder_oms/cuda/CUDA_OMS_SIMD.cu:115-116); C if none
 * early_term: 1 = per-word stop after a non-first iteration whose checks all
 *            have even parity of positive v2c values
 *            (src/decoder_2nms/cuda/CUDA_2NMS_SIMD.cu:164-166,237-244)
 * stopped  : [W] per-word stop flags (in/out), may be NULL when !early_term
 * table    : learned table [iter_base + it][degree] (LOMS: offset in LSB,
 *            LNMS: numerator over 8, result capped at 31), row length tdeg,
 *            rows tits; else NULL
 * iter_base: absolute index of this call's first iteration (learned tables)
 * iters_run: [W] out, iterations executed per word in this call (may be NULL)
 * returns 0, or -1 on bad arguments. */
int orc_decode(int N, int C, const int *deg, const uint32_t *pos, int variant,
               size_t W, uint32_t *post, uint32_t *msg, int iters, int first,
               int q1_split, int early_term, uint8_t *stopped,
               const int8_t *table, int tits, int tdeg, int iter_base,
               int *iters_run);

/* Same, parallelised over words with `threads` POSIX threads (CPU baseline). */
int orc_decode_mt(int N, int C, const int *deg, const uint32_t *pos, int variant,
                  size_t W, uint32_t *post, uint32_t *msg, int iters, int first,
                  int q1_split, int early_term, uint8_t *stopped,
                  const int8_t *table, int tits, int tdeg, int iter_base,
                  int *iters_run, int threads);

/* Per-frame stop (the persistent frame-queue decoder's model): each frame is
 * decoded alone (in its own byte lane) from iteration 0 and stops after the
 * first iteration (the first included, up to max_iters) at which
 *   stop_rule 0: its hard decisions (post > 0) satisfy every check, or
 *   stop_rule 1: every check had an even parity of positive v2c values
 *                (the test of CUDA_2NMS_SIMD.cu:237-244).
 * llr / hard: frame-major [F][N]; iters[F] (may be NULL) = iterations run.
 * Variant OMS is refused: its first-iteration rule depends on the degree
 * group (use OMS_V2 = OMS on a code with one group). */
int orc_decode_frames(int N, int C, const int *deg, const uint32_t *pos, int variant,
                      const int8_t *llr, size_t F, int max_iters, const int8_t *table,
                      int tits, int tdeg, uint8_t *hard, int *iters, int stop_rule, int threads);

/* ---- layouts (src/transpose/GPU_Transpose_uint8.cu:9-130) --------------- */
/* frame-major int8 in[(4w+k)*N + v]  ->  post[v*W + w] byte k               */
void orc_interleave(const int8_t *in, uint32_t *post, int N, size_t W);
/* post -> frame-major hard decision out[(4w+k)*N + v] = (byte > 0)          */
void orc_deinterleave_hd(const uint32_t *post, uint8_t *out, int N, size_t W);

/* ---- channel (src/awgn_channel/CChanel_AWGN_SIMD.cu:7-30, 63-73) --------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* sigma for BPSK at Eb/N0 (dB) and code rate R (CChanel_AWGN_SIMD.cu:63-73)  */
double orc_sigma(double ebn0_db, double rate);
/* Standard normal sample number `s` of stream (seed, snr_idx).              */
double orc_gauss(uint64_t seed, uint32_t snr_idx, uint64_t s);
/* Quantised LLR of all-zero codeword sample: BPSK bit0 -> -1, y = -1+sigma*z.
 * quant 0: sat127(round(8*2y/sigma^2))  quant 1: sat31(trunc(8*y)) (reference) */
int8_t orc_llr(uint64_t seed, uint32_t snr_idx, uint64_t s, double sigma, int quant);
/* Fill frame-major LLRs for frames [f0, f0+nf) of length N.                  */
void orc_channel_frames(int8_t *out, int N, size_t f0, size_t nf, uint64_t seed,
                        uint32_t snr_idx, double sigma, int quant);

/* Frames [f0, f0+nf) with per-frame Eb/N0 = lo + (hi-lo)*u(f), u from the
 * Philox block (f_lo, f_hi, snr_idx, 0x534E5252), sigma = sqrt(0.5/(R 10^(EbN0/10))). */
void orc_channel_frames_range(int8_t *out, int N, size_t f0, size_t nf, uint64_t seed,
                              uint32_t snr_idx, double lo_db, double hi_db, double rate, int quant);

/* ---- error counting (src/ber_analyzer/CErrorAnalyzer.cpp:141-157) -------- */
/* hard[f*N + v]; counts positions v < n_info. Returns frames in error.       */
long long orc_count_errors(const uint8_t *hard, int N, size_t frames, int n_info,
                           long long *bit_errors, uint32_t *per_frame_bits);
/* Per frame syndrome check: 1 = all parity checks satisfied.                */
void orc_syndrome(const uint8_t *hard, int N, int C, const int *deg,
                  const uint32_t *pos, size_t frames, uint8_t *ok);

/* ---- float shadow of the layered decoder (learned tables) --------------
 * Same serial table order as orc_decode, in float LSB units:
 *   t = clamp(p - m, -128, 127)   (first iteration: m = 0)
 *   mo_j = min of the OTHER |t| (127 cap), sign by parity of the others' t > 0
 *   mode 0 (offset): mag = clamp(mo - beta[it][d], 0, 31)
 *   mode 1 (scale):  mag = clamp(alpha[it][d] * mo, 0, 31)
 *   m_j = +-mag, p = clamp(t + m_j, -128, 127)
 * loss per frame = mean over iterations and variables of softplus(p / 8)
 * (all-zero codeword: a positive value votes for bit 1).
 * params / grad: [T][dcols], column = min(degree, dcols - 1).
 * Returns the mean loss over F frames; grad (may be NULL) is its gradient. */
double orc_shadow_loss_grad(int N, int C, const int *deg, const uint32_t *pos, int mode, int T,
                            int dcols, const float *params, const int8_t *llr, size_t F,
                            float *grad);
/* Same forward pass; writes each frame's posterior after T iterations
 * (post_out[f*N + v]) and returns the mean loss. */
double orc_shadow_forward(int N, int C, const int *deg, const uint32_t *pos, int mode, int T,
                          int dcols, const float *params, const int8_t *llr, size_t F,
                          float *post_out);

#ifdef __cplusplus
}
#endif
#endif
