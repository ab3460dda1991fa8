/*
 * ldpc_oracle.c -- CPU restatement of the reference's integer Min-Sum path.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py as the checker; the shipped HIP library
 * never links or calls it.
 *
 * Structure deliberately follows the reference's per-thread kernel: one
 * uint32_t word carries 4 frames (one int8 lane each) and the checks are
 * processed serially in table order, posterior updated in place after every
 * check.  The per-byte operations are written from their mathematical
 * definitions (src/decoder_ms/cuda/simd_functions.h:89-131), byte by byte
 * with plain int arithmetic.
 */
#include "ldpc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* per-byte ops                                                              */
/* ------------------------------------------------------------------------ */
static inline int sb(uint32_t x, int k) { return (int8_t)((x >> (8 * k)) & 0xFFu); }
static inline int ub(uint32_t x, int k) { return (int)((x >> (8 * k)) & 0xFFu); }
static inline uint32_t put(uint32_t acc, int k, int v) { return acc | ((uint32_t)(v & 0xFF) << (8 * k)); }
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* vaddss4: per-byte signed saturating add (simd_functions.h:96) */
uint32_t orc_vaddss4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) r = put(r, k, clampi(sb(a, k) + sb(b, k), -128, 127));
    return r;
}
/* vsubss4: per-byte signed saturating subtract (simd_functions.h:129) */
uint32_t orc_vsubss4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) r = put(r, k, clampi(sb(a, k) - sb(b, k), -128, 127));
    return r;
}
/* vsubus4: per-byte unsigned saturating subtract (simd_functions.h:130) */
uint32_t orc_vsubus4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) r = put(r, k, clampi(ub(a, k) - ub(b, k), 0, 255));
    return r;
}
/* vminu4 / vmaxu4: per-byte unsigned min / max (simd_functions.h:111,109) */
uint32_t orc_vminu4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) { int x = ub(a, k), y = ub(b, k); r = put(r, k, x < y ? x : y); }
    return r;
}
uint32_t orc_vmaxu4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) { int x = ub(a, k), y = ub(b, k); r = put(r, k, x > y ? x : y); }
    return r;
}
/* vabs4: per-byte |a| with wrap-around, |-128| = 0x80 (simd_functions.h:91) */
uint32_t orc_vabs4(uint32_t a) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) { int x = sb(a, k); r = put(r, k, x < 0 ? -x : x); }
    return r;
}
/* vneg4: per-byte wrap-around negation (simd_functions.h:112) */
uint32_t orc_vneg4(uint32_t a) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) r = put(r, k, -sb(a, k));
    return r;
}
/* vcmpgts4: per-byte signed a > b ? 0xff : 0 (simd_functions.h:102) */
uint32_t orc_vcmpgts4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) r = put(r, k, sb(a, k) > sb(b, k) ? 0xFF : 0);
    return r;
}
/* vcmpeq4 / vcmpne4: per-byte equality masks (simd_functions.h:99,108) */
uint32_t orc_vcmpeq4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) r = put(r, k, ub(a, k) == ub(b, k) ? 0xFF : 0);
    return r;
}
uint32_t orc_vcmpne4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) r = put(r, k, ub(a, k) != ub(b, k) ? 0xFF : 0);
    return r;
}
/* vsetgts4: per-byte signed a > b ? 1 : 0 (simd_functions.h:122) */
uint32_t orc_vsetgts4(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) r = put(r, k, sb(a, k) > sb(b, k) ? 1 : 0);
    return r;
}

/* ------------------------------------------------------------------------ */
/* magnitude functions per variant                                           */
/* ------------------------------------------------------------------------ */
/* NMS / 2NMS scale each byte through a float multiply and truncation to char:
 * src/decoder_nms/cuda/CUDA_NMS_SIMD.cu:73-85 (0.75),
 * src/decoder_2nms/cuda/CUDA_2NMS_SIMD.cu:80-83 (0.875 on min2). */
static uint32_t scale_bytes(uint32_t m, float f) {
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) {
        signed char c = (signed char)sb(m, k);
        signed char s = (signed char)(((float)c) * f);
        r = put(r, k, s);
    }
    return r;
}

/* cste_1 (value given to the min edge, from min2) and cste_2 (others, from
 * min1) for one check.  `nocap` = OMS first-iteration second degree group
 * (src/decoder_oms/cuda/CUDA_OMS_SIMD.cu:115-116). */
static void constants(int variant, uint32_t min1, uint32_t min2, int nocap,
                      int lparam, uint32_t *c1, uint32_t *c2) {
    switch (variant) {
    case ORC_MS: /* CUDA_MS_SIMD.cu:73-74 */
        *c1 = orc_vminu4(min2, 0x1F1F1F1Fu);
        *c2 = orc_vminu4(min1, 0x1F1F1F1Fu);
        break;
    case ORC_OMS: /* CUDA_OMS_SIMD.cu:73-74, 115-116 */
        if (nocap) {
            *c1 = orc_vsubus4(min2, 0x01010101u);
            *c2 = orc_vsubus4(min1, 0x01010101u);
        } else {
            *c1 = orc_vminu4(orc_vsubus4(min2, 0x01010101u), 0x1F1F1F1Fu);
            *c2 = orc_vminu4(orc_vsubus4(min1, 0x01010101u), 0x1F1F1F1Fu);
        }
        break;
    case ORC_OMS_V2: /* CUDA_OMS_SIMD_v2.cu:82-83, 145-146 (clamp everywhere) */
        *c1 = orc_vminu4(orc_vsubus4(min2, 0x01010101u), 0x1F1F1F1Fu);
        *c2 = orc_vminu4(orc_vsubus4(min1, 0x01010101u), 0x1F1F1F1Fu);
        break;
    case ORC_NMS: /* CUDA_NMS_SIMD.cu:73-85 */
        *c1 = scale_bytes(min2, 0.750f);
        *c2 = scale_bytes(min1, 0.750f);
        break;
    case ORC_2NMS: /* CUDA_2NMS_SIMD.cu:73-85 */
        *c1 = scale_bytes(min2, 0.875f);
        *c2 = scale_bytes(min1, 0.750f);
        break;
    case ORC_LOMS: { /* learned offset: min(sat_u(m - beta), 31) */
        uint32_t b4 = (uint32_t)(lparam & 0xFF) * 0x01010101u;
        *c1 = orc_vminu4(orc_vsubus4(min2, b4), 0x1F1F1F1Fu);
        *c2 = orc_vminu4(orc_vsubus4(min1, b4), 0x1F1F1F1Fu);
        break;
    }
    case ORC_LNMS: { /* learned normalisation: min(floor(m * a / 8), 31), a in [0,8] */
        uint32_t r1 = 0, r2 = 0;
        for (int k = 0; k < 4; k++) {
            int x1 = (ub(min2, k) * lparam) >> 3, x2 = (ub(min1, k) * lparam) >> 3;
            r1 = put(r1, k, x1 < 31 ? x1 : 31);
            r2 = put(r2, k, x2 < 31 ? x2 : 31);
        }
        *c1 = r1;
        *c2 = r2;
        break;
    }
    default:
        *c1 = *c2 = 0;
    }
}

/* ------------------------------------------------------------------------ */
/* one check on one word: CUDA_OMS_SIMD.cu:148-191 (iterations >= 2) and     */
/* :47-88 (first iteration, no message read)                                 */
/* ------------------------------------------------------------------------ */
#define ORC_MAXD 64
static uint32_t check_word(int d, const uint32_t *vars, uint32_t *post, uint32_t *msg,
                           size_t W, size_t w, int first, int variant, int nocap,
                           int lparam) {
    uint32_t t[ORC_MAXD];
    uint32_t sign = 0, min1 = 0x7F7F7F7Fu, min2 = 0x7F7F7F7Fu;
    for (int j = 0; j < d; j++) {
        uint32_t p = post[(size_t)vars[j] * W + w];
        t[j] = first ? p : orc_vsubss4(p, msg[(size_t)j * W + w]);      /* :166 */
        uint32_t a = orc_vabs4(t[j]);
        min2 = orc_vminu4(min2, orc_vmaxu4(a, min1));                    /* :168 */
        min1 = orc_vminu4(min1, a);                                      /* :169 */
        sign ^= orc_vcmpgts4(t[j], 0);                                   /* :170 */
    }
    uint32_t c1, c2;
    constants(variant, min1, min2, nocap, lparam, &c1, &c2);             /* :173-174 */
    for (int j = 0; j < d; j++) {
        uint32_t ab = orc_vabs4(t[j]);
        uint32_t m1 = orc_vcmpeq4(ab, min1);
        uint32_t m2 = orc_vcmpne4(ab, min1);
        uint32_t re = (m1 & c1) | (m2 & c2);                             /* :181 */
        uint32_t sm = sign ^ orc_vcmpgts4(t[j], 0);                      /* :182 */
        uint32_t out = (re & sm) | (orc_vneg4(re) & ~sm);                /* :183 */
        msg[(size_t)j * W + w] = out;                                    /* :184 */
        post[(size_t)vars[j] * W + w] = orc_vaddss4(t[j], out);          /* :186 */
    }
    return sign;
}

static int table_param(int variant, const int8_t *table, int tits, int tdeg,
                       int abs_it, int d) {
    if (variant != ORC_LOMS && variant != ORC_LNMS) return 0;
    int row = abs_it < tits ? abs_it : tits - 1;
    int col = d < tdeg ? d : tdeg - 1;
    return table[(size_t)row * tdeg + col];
}

/* Decode words [w0, w1). */
static void decode_range(int N, int C, const int *deg, const uint32_t *pos,
                         const size_t *eoff, int variant, size_t W, uint32_t *post,
                         uint32_t *msg, int iters, int first, int q1_split,
                         int early_term, uint8_t *stopped, const int8_t *table,
                         int tits, int tdeg, int iter_base, int *iters_run,
                         size_t w0, size_t w1) {
    (void)N;
    for (size_t w = w0; w < w1; w++) {
        int ran = 0;
        if (stopped && stopped[w]) { if (iters_run) iters_run[w] = 0; continue; }
        for (int it = 0; it < iters; it++) {
            int is_first = first && it == 0;
            uint32_t ov = 0;
            for (int c = 0; c < C; c++) {
                int d = deg[c];
                int nocap = is_first && variant == ORC_OMS && c >= q1_split;
                int lp = table_param(variant, table, tits, tdeg, iter_base + it, d);
                ov |= check_word(d, pos + eoff[c], post, msg + eoff[c] * W, W, w,
                                 is_first, variant, nocap, lp);
            }
            ran++;
            /* CUDA_2NMS_SIMD.cu:237-244: stop test after every iteration of the
             * while loop, i.e. every iteration except a decode's first one. */
            if (early_term && !is_first && ov == 0) {
                if (stopped) stopped[w] = 1;
                break;
            }
        }
        if (iters_run) iters_run[w] = ran;
    }
}

static int validate(int N, int C, const int *deg, const uint32_t *pos, int variant,
                    int iters, const int8_t *table, int tits, int tdeg, size_t **eoff_out) {
    if (N <= 0 || C <= 0 || !deg || !pos || iters < 1) return -1;
    if (variant < ORC_MS || variant > ORC_LNMS) return -1;
    if ((variant == ORC_LOMS || variant == ORC_LNMS) && (!table || tits < 1 || tdeg < 1)) return -1;
    size_t *eoff = (size_t *)malloc(sizeof(size_t) * (size_t)(C + 1));
    if (!eoff) return -1;
    eoff[0] = 0;
    for (int c = 0; c < C; c++) {
        if (deg[c] < 1 || deg[c] > ORC_MAXD) { free(eoff); return -1; }
        eoff[c + 1] = eoff[c] + (size_t)deg[c];
    }
    for (size_t e = 0; e < eoff[C]; e++)
        if (pos[e] >= (uint32_t)N) { free(eoff); return -1; }
    *eoff_out = eoff;
    return 0;
}

int orc_decode(int N, int C, const int *deg, const uint32_t *pos, int variant,
               size_t W, uint32_t *post, uint32_t *msg, int iters, int first,
               int q1_split, int early_term, uint8_t *stopped,
               const int8_t *table, int tits, int tdeg, int iter_base,
               int *iters_run) {
    size_t *eoff;
    if (validate(N, C, deg, pos, variant, iters, table, tits, tdeg, &eoff)) return -1;
    decode_range(N, C, deg, pos, eoff, variant, W, post, msg, iters, first, q1_split,
                 early_term, stopped, table, tits, tdeg, iter_base, iters_run, 0, W);
    free(eoff);
    return 0;
}

typedef struct {
    int N, C; const int *deg; const uint32_t *pos; const size_t *eoff; int variant;
    size_t W; uint32_t *post; uint32_t *msg; int iters, first, q1_split, early_term;
    uint8_t *stopped; const int8_t *table; int tits, tdeg, iter_base; int *iters_run;
    size_t w0, w1;
} job_t;

static void *job_run(void *p) {
    job_t *j = (job_t *)p;
    decode_range(j->N, j->C, j->deg, j->pos, j->eoff, j->variant, j->W, j->post, j->msg,
                 j->iters, j->first, j->q1_split, j->early_term, j->stopped, j->table,
                 j->tits, j->tdeg, j->iter_base, j->iters_run, j->w0, j->w1);
    return NULL;
}

int orc_decode_mt(int N, int C, const int *deg, const uint32_t *pos, int variant,
                  size_t W, uint32_t *post, uint32_t *msg, int iters, int first,
                  int q1_split, int early_term, uint8_t *stopped,
                  const int8_t *table, int tits, int tdeg, int iter_base,
                  int *iters_run, int threads) {
    size_t *eoff;
    if (validate(N, C, deg, pos, variant, iters, table, tits, tdeg, &eoff)) return -1;
    if (threads < 1) threads = 1;
    if ((size_t)threads > W) threads = (int)(W ? W : 1);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    job_t *jobs = (job_t *)malloc(sizeof(job_t) * (size_t)threads);
    for (int t = 0; t < threads; t++) {
        job_t j = {N, C, deg, pos, eoff, variant, W, post, msg, iters, first, q1_split,
                   early_term, stopped, table, tits, tdeg, iter_base, iters_run,
                   W * (size_t)t / (size_t)threads, W * (size_t)(t + 1) / (size_t)threads};
        jobs[t] = j;
        pthread_create(&th[t], NULL, job_run, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    free(eoff);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* per-frame stop, two rules, tested after every iteration (the first     */
/* included) on each byte lane on its own:                                  */
/*   rule 0 (syndrome): the hard decisions (post > 0, GPU_Transpose_uint8  */
/*     .cu:29) satisfy every check;                                         */
/*   rule 1 (v2c parity): every check had an even parity of positive v2c   */
/*     values during the iteration (CUDA_2NMS_SIMD.cu:237-244).             */
/* Lanes never mix (every operation above is per byte), so four frames     */
/* share a word; a frame's result is its decisions and iteration count when */
/* its lane stops.                                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    int N, C; const int *deg; const uint32_t *pos; const size_t *eoff; int variant;
    const int8_t *llr; size_t F; int max_iters; const int8_t *table; int tits, tdeg;
    uint8_t *hard; int *iters; size_t g0, g1; int stop_rule;
} fjob_t;

/* per byte: parity of (post > 0) over the check's variables, 0xFF/0 masks */
static uint32_t syndrome_word(int d, const uint32_t *vars, const uint32_t *post) {
    uint32_t s = 0;
    for (int j = 0; j < d; j++) s ^= orc_vcmpgts4(post[vars[j]], 0);
    return s;
}

static void frames_range(const fjob_t *j) {
    const int N = j->N;
    uint32_t *post = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)N);
    uint32_t *msg = (uint32_t *)malloc(sizeof(uint32_t) * (j->eoff[j->C] ? j->eoff[j->C] : 1));
    for (size_t g = j->g0; g < j->g1; g++) {
        size_t f0 = 4 * g;
        int nf = (int)((j->F - f0) < 4 ? (j->F - f0) : 4), done = 0;
        for (int v = 0; v < N; v++) {
            uint32_t x = 0;
            for (int k = 0; k < nf; k++) x = put(x, k, j->llr[(f0 + k) * (size_t)N + v]);
            post[v] = x;
        }
        for (int k = nf; k < 4; k++) done |= 1 << k;
        for (int it = 0; it < j->max_iters && done != 15; it++) {
            uint32_t ov = 0;
            for (int c = 0; c < j->C; c++) {
                int d = j->deg[c];
                int lp = table_param(j->variant, j->table, j->tits, j->tdeg, it, d);
                ov |= check_word(d, j->pos + j->eoff[c], post, msg + j->eoff[c], 1, 0, it == 0,
                                 j->variant, 0, lp);
            }
            if (j->stop_rule == 0) {
                ov = 0;
                for (int c = 0; c < j->C; c++) ov |= syndrome_word(j->deg[c], j->pos + j->eoff[c], post);
            }
            for (int k = 0; k < 4; k++) {
                if (done >> k & 1) continue;
                if (ub(ov, k) == 0 || it + 1 == j->max_iters) {
                    uint8_t *h = j->hard + (f0 + k) * (size_t)N;
                    for (int v = 0; v < N; v++) h[v] = (uint8_t)(sb(post[v], k) > 0);
                    if (j->iters) j->iters[f0 + k] = it + 1;
                    done |= 1 << k;
                }
            }
        }
    }
    free(post);
    free(msg);
}

static void *fjob_run(void *p) { frames_range((const fjob_t *)p); return NULL; }

int orc_decode_frames(int N, int C, const int *deg, const uint32_t *pos, int variant,
                      const int8_t *llr, size_t F, int max_iters, const int8_t *table,
                      int tits, int tdeg, uint8_t *hard, int *iters, int stop_rule, int threads) {
    size_t *eoff;
    if (!llr || !hard || variant == ORC_OMS || (stop_rule != 0 && stop_rule != 1)) return -1;
    if (validate(N, C, deg, pos, variant, max_iters, table, tits, tdeg, &eoff)) return -1;
    size_t G = (F + 3) / 4;
    if (threads < 1) threads = 1;
    if ((size_t)threads > G) threads = (int)(G ? G : 1);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    fjob_t *jobs = (fjob_t *)malloc(sizeof(fjob_t) * (size_t)threads);
    for (int t = 0; t < threads; t++) {
        fjob_t j = {N, C, deg, pos, eoff, variant, llr, F, max_iters, table, tits, tdeg, hard, iters,
                    G * (size_t)t / (size_t)threads, G * (size_t)(t + 1) / (size_t)threads, stop_rule};
        jobs[t] = j;
        pthread_create(&th[t], NULL, fjob_run, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    free(eoff);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* layouts: closed form of src/transpose/GPU_Transpose_uint8.cu:80-130       */
/* (Interleaver_uint8) and :9-78 (InvInterleaver_uint8 + vsetgts4 at :29)    */
/* ------------------------------------------------------------------------ */
void orc_interleave(const int8_t *in, uint32_t *post, int N, size_t W) {
    for (int v = 0; v < N; v++)
        for (size_t w = 0; w < W; w++) {
            uint32_t x = 0;
            for (int k = 0; k < 4; k++) x = put(x, k, in[(4 * w + (size_t)k) * (size_t)N + (size_t)v]);
            post[(size_t)v * W + w] = x;
        }
}

void orc_deinterleave_hd(const uint32_t *post, uint8_t *out, int N, size_t W) {
    for (int v = 0; v < N; v++)
        for (size_t w = 0; w < W; w++) {
            uint32_t hd = orc_vsetgts4(post[(size_t)v * W + w], 0);
            for (int k = 0; k < 4; k++)
                out[(4 * w + (size_t)k) * (size_t)N + (size_t)v] = (uint8_t)ub(hd, k);
        }
}

/* ------------------------------------------------------------------------ */
/* channel                                                                   */
/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
 * as easy as 1, 2, 3"): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85, ten rounds.                            */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* src/awgn_channel/CChanel_AWGN_SIMD.cu:63-73: sigma = sqrt(10^(-(EbN0+10log10 R)/10)/2) */
double orc_sigma(double ebn0_db, double rate) {
    double interm = -0.1 * (ebn0_db + 10.0 * log10(rate));
    return sqrt(pow(10.0, interm) / 2.0);
}

/* Box-Muller (cosine branch) on two 53-bit uniforms from one Philox block:
 * counter = (s_lo, s_hi, snr_idx, 0x4C445043), key = (seed_lo, seed_hi).   */
double orc_gauss(uint64_t seed, uint32_t snr_idx, uint64_t s) {
    uint32_t ctr[4] = {(uint32_t)s, (uint32_t)(s >> 32), snr_idx, 0x4C445043u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    uint64_t a = (((uint64_t)r[0] << 32) | r[1]) >> 11;
    uint64_t b = (((uint64_t)r[2] << 32) | r[3]) >> 11;
    double u1 = ((double)a + 1.0) * 0x1.0p-53; /* (0, 1] */
    double u2 = (double)b * 0x1.0p-53;         /* [0, 1) */
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

int8_t orc_llr(uint64_t seed, uint32_t snr_idx, uint64_t s, double sigma, int quant) {
    double y = -1.0 + sigma * orc_gauss(seed, snr_idx, s);
    if (quant == 1) {
        /* CChanel_AWGN_SIMD.cu:21-24: (char)fminf(fmaxf(8*v, -31), 31) */
        float v = (float)y;
        float q = fminf(fmaxf(8.0f * v, -31.0f), 31.0f);
        return (int8_t)q;
    }
    const double scale = 16.0 / (sigma * sigma); /* 8 * (2y / sigma^2) */
    double q = round(scale * y);
    if (q > 127.0) q = 127.0;
    if (q < -127.0) q = -127.0;
    return (int8_t)q;
}

void orc_channel_frames(int8_t *out, int N, size_t f0, size_t nf, uint64_t seed,
                        uint32_t snr_idx, double sigma, int quant) {
    for (size_t f = 0; f < nf; f++)
        for (int v = 0; v < N; v++)
            out[f * (size_t)N + (size_t)v] =
                orc_llr(seed, snr_idx, (f0 + f) * (uint64_t)N + (uint64_t)v, sigma, quant);
}

void orc_channel_frames_range(int8_t *out, int N, size_t f0, size_t nf, uint64_t seed,
                              uint32_t snr_idx, double lo_db, double hi_db, double rate, int quant) {
    for (size_t f = 0; f < nf; f++) {
        const uint64_t fi = f0 + f;
        uint32_t ctr[4] = {(uint32_t)fi, (uint32_t)(fi >> 32), snr_idx, 0x534E5252u};
        uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        uint32_t r[4];
        orc_philox4x32_10(ctr, key, r);
        const double ebn0 = lo_db + (hi_db - lo_db) * ((double)r[0] * 0x1.0p-32);
        const double sigma = sqrt(0.5 / (rate * pow(10.0, ebn0 / 10.0)));
        for (int v = 0; v < N; v++)
            out[f * (size_t)N + (size_t)v] = orc_llr(seed, snr_idx, fi * (uint64_t)N + (uint64_t)v, sigma, quant);
    }
}

/* ------------------------------------------------------------------------ */
/* error counting: src/ber_analyzer/CErrorAnalyzer.cpp:141-157 (counts the   */
/* first _vars = N - K positions of each frame, all-zero codeword)           */
/* ------------------------------------------------------------------------ */
long long orc_count_errors(const uint8_t *hard, int N, size_t frames, int n_info,
                           long long *bit_errors, uint32_t *per_frame_bits) {
    long long fe = 0, be = 0;
    for (size_t f = 0; f < frames; f++) {
        uint32_t e = 0;
        for (int v = 0; v < n_info; v++) e += hard[f * (size_t)N + (size_t)v] != 0;
        be += e;
        fe += e != 0;
        if (per_frame_bits) per_frame_bits[f] = e;
    }
    if (bit_errors) *bit_errors = be;
    return fe;
}

void orc_syndrome(const uint8_t *hard, int N, int C, const int *deg,
                  const uint32_t *pos, size_t frames, uint8_t *ok) {
    for (size_t f = 0; f < frames; f++) {
        const uint8_t *h = hard + f * (size_t)N;
        size_t e = 0;
        int good = 1;
        for (int c = 0; c < C; c++) {
            int par = 0;
            for (int j = 0; j < deg[c]; j++) par ^= h[pos[e + (size_t)j]] & 1;
            e += (size_t)deg[c];
            if (par) good = 0;
        }
        ok[f] = (uint8_t)good;
    }
}

/* ------------------------------------------------------------------------ */
/* float shadow decoder: forward (stores pre-clamp v2c values r and the      */
/* posterior after each iteration) and reverse-mode gradient w.r.t. the      */
/* per-iteration, per-degree table.  Reverse sweep: checks in reverse table   */
/* order; the gradient reaching a message goes to the edge holding min1       */
/* (or min2 for the minimum edge itself).                                     */
/* ------------------------------------------------------------------------ */
static float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
static float softplusf(float x) { return (x > 0 ? x : 0.0f) + log1pf(expf(-fabsf(x))); }
static float sigmoidf(float x) { return 1.0f / (1.0f + expf(-x)); }

/* recompute one check from its pre-clamp values r[0..d) */
static void shadow_check(int d, const float *r, float par_val, int mode, float *t, float *mnew,
                         int *i1o, int *i2o, float *min1o, float *min2o) {
    float min1 = 127.0f, min2 = 127.0f;
    int i1 = -1, i2 = -1, par = 0;
    for (int j = 0; j < d; j++) {
        t[j] = clampf(r[j], -128.0f, 127.0f);
        float a = fabsf(t[j]);
        if (a < min1) { min2 = min1; i2 = i1; min1 = a; i1 = j; }
        else if (a < min2) { min2 = a; i2 = j; }
        par ^= t[j] > 0.0f;
    }
    for (int j = 0; j < d; j++) {
        float mo = (j == i1) ? min2 : min1;
        float mag = mode == 0 ? clampf(mo - par_val, 0.0f, 31.0f) : clampf(par_val * mo, 0.0f, 31.0f);
        int pos_others = par ^ (t[j] > 0.0f);
        mnew[j] = pos_others ? mag : -mag;
    }
    *i1o = i1; *i2o = i2; *min1o = min1; *min2o = min2;
}

static double shadow_run(int N, int C, const int *deg, const uint32_t *pos, int mode, int T,
                         int dcols, const float *params, const int8_t *llr, size_t F, float *grad,
                         float *post_out);

double orc_shadow_loss_grad(int N, int C, const int *deg, const uint32_t *pos, int mode, int T,
                            int dcols, const float *params, const int8_t *llr, size_t F,
                            float *grad) {
    return shadow_run(N, C, deg, pos, mode, T, dcols, params, llr, F, grad, NULL);
}

double orc_shadow_forward(int N, int C, const int *deg, const uint32_t *pos, int mode, int T,
                          int dcols, const float *params, const int8_t *llr, size_t F,
                          float *post_out) {
    return shadow_run(N, C, deg, pos, mode, T, dcols, params, llr, F, NULL, post_out);
}

static double shadow_run(int N, int C, const int *deg, const uint32_t *pos, int mode, int T,
                         int dcols, const float *params, const int8_t *llr, size_t F, float *grad,
                         float *post_out) {
    size_t *eoff = (size_t *)malloc(sizeof(size_t) * (size_t)(C + 1));
    eoff[0] = 0;
    for (int c = 0; c < C; c++) eoff[c + 1] = eoff[c] + (size_t)deg[c];
    const size_t E = eoff[C];
    float *p = (float *)malloc(sizeof(float) * (size_t)N);
    float *m = (float *)malloc(sizeof(float) * E);
    float *r = (float *)malloc(sizeof(float) * E * (size_t)T);
    float *pend = (float *)malloc(sizeof(float) * (size_t)N * (size_t)T);
    float *gp = (float *)malloc(sizeof(float) * (size_t)N);
    float *gm = (float *)malloc(sizeof(float) * E);
    if (grad) memset(grad, 0, sizeof(float) * (size_t)T * (size_t)dcols);
    double loss = 0.0;
    float t[ORC_MAXD], mn[ORC_MAXD], ga[ORC_MAXD], gt[ORC_MAXD];
    const float wl = 1.0f / ((float)T * (float)N);
    for (size_t f = 0; f < F; f++) {
        for (int v = 0; v < N; v++) p[v] = (float)llr[f * (size_t)N + (size_t)v];
        memset(m, 0, sizeof(float) * E);
        double lf = 0.0;
        for (int it = 0; it < T; it++) {
            float *ri = r + (size_t)it * E;
            for (int c = 0; c < C; c++) {
                const int d = deg[c];
                const size_t e0 = eoff[c];
                const float pv = params[(size_t)it * dcols + (d < dcols ? d : dcols - 1)];
                for (int j = 0; j < d; j++) ri[e0 + j] = p[pos[e0 + j]] - (it == 0 ? 0.0f : m[e0 + j]);
                int i1, i2;
                float mi1, mi2;
                shadow_check(d, ri + e0, pv, mode, t, mn, &i1, &i2, &mi1, &mi2);
                for (int j = 0; j < d; j++) {
                    m[e0 + j] = mn[j];
                    p[pos[e0 + j]] = clampf(t[j] + mn[j], -128.0f, 127.0f);
                }
            }
            for (int v = 0; v < N; v++) {
                pend[(size_t)it * N + v] = p[v];
                lf += softplusf(p[v] / 8.0f);
            }
        }
        loss += lf * wl;
        if (post_out) memcpy(post_out + f * (size_t)N, p, sizeof(float) * (size_t)N);
        if (!grad) continue;
        memset(gp, 0, sizeof(float) * (size_t)N);
        memset(gm, 0, sizeof(float) * E);
        for (int it = T - 1; it >= 0; it--) {
            for (int v = 0; v < N; v++) gp[v] += wl * sigmoidf(pend[(size_t)it * N + v] / 8.0f) / 8.0f;
            const float *ri = r + (size_t)it * E;
            for (int c = C - 1; c >= 0; c--) {
                const int d = deg[c];
                const size_t e0 = eoff[c];
                const int col = d < dcols ? d : dcols - 1;
                const float pv = params[(size_t)it * dcols + col];
                int i1, i2;
                float mi1, mi2;
                shadow_check(d, ri + e0, pv, mode, t, mn, &i1, &i2, &mi1, &mi2);
                int par = 0;
                for (int k = 0; k < d; k++) par ^= t[k] > 0.0f;
                for (int j = 0; j < d; j++) ga[j] = 0.0f;
                for (int j = 0; j < d; j++) {
                    /* p_j' = clamp(t_j + m_j): gradient passes inside the clamp */
                    const float u = t[j] + mn[j];
                    const float gu = (u > -128.0f && u < 127.0f) ? gp[pos[e0 + j]] : 0.0f;
                    gt[j] = gu;
                    /* m_j feeds p_j' and, through gm, the next iteration's t_j */
                    const float gmn = gu + gm[e0 + j];
                    const float s = (par ^ (t[j] > 0.0f)) ? 1.0f : -1.0f;
                    const float gmag = s * gmn;
                    const float mo = (j == i1) ? mi2 : mi1;
                    float gmo;
                    if (mode == 0) {
                        const float x = mo - pv;
                        const int inside = x > 0.0f && x < 31.0f;
                        gmo = inside ? gmag : 0.0f;
                        grad[(size_t)it * dcols + col] -= inside ? gmag : 0.0f;
                    } else {
                        const float x = pv * mo;
                        const int inside = x > 0.0f && x < 31.0f;
                        gmo = inside ? gmag * pv : 0.0f;
                        grad[(size_t)it * dcols + col] += inside ? gmag * mo : 0.0f;
                    }
                    /* mo is |t| of the edge holding min1, or min2 for that edge itself */
                    if (j == i1) {
                        if (i2 >= 0) ga[i2] += gmo;
                    } else if (i1 >= 0) {
                        ga[i1] += gmo;
                    }
                }
                for (int k = 0; k < d; k++) {
                    const float sg = t[k] > 0.0f ? 1.0f : (t[k] < 0.0f ? -1.0f : 0.0f);
                    gt[k] += ga[k] * sg;
                }
                for (int j = 0; j < d; j++) {
                    const float rj = ri[e0 + j];
                    const int inside = rj > -128.0f && rj < 127.0f;
                    gp[pos[e0 + j]] = inside ? gt[j] : 0.0f;
                    gm[e0 + j] = (it > 0 && inside) ? -gt[j] : 0.0f;
                }
            }
        }
    }
    if (grad)
        for (int i = 0; i < T * dcols; i++) grad[i] /= (float)F;
    free(p); free(m); free(r); free(pend); free(gp); free(gm); free(eoff);
    return F ? loss / (double)F : 0.0;
}
