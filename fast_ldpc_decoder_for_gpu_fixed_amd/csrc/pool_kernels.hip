// pool_kernels.hip -- persistent decoding of a frame queue with per-frame
// early stop and slot refill.
//
// A workgroup keeps 4P frame slots resident: P words of 4 byte lanes, the
// posterior of all of them in LDS ([var][P] words, as in k_decode).  It runs
// the layered iteration on all slots at once; after each iteration a frame
// that passes its stop test or reached the iteration limit writes its
// decisions and iteration count, and its lane is refilled with the next
// frame of the queue (one atomic per workgroup and iteration).  Stop tests:
//   syndrome   -- the decisions (post > 0) satisfy every check: one read-only
//                 pass over all checks after the iteration;
//   v2c parity -- every check had an even parity of positive v2c values
//                 during the iteration (the reference's rule,
//                 src/decoder_2nms/cuda/CUDA_2NMS_SIMD.cu:237-244, per frame).
// A
// fresh lane reads zero messages in its first iteration: the per-word byte
// mask `keep` clears that lane of every message word it loads.  Learned
// tables are looked up per lane with the lane's own iteration index.
// The workgroup exits when the queue is empty and every slot is drained.
#include "check_core.h"

namespace ldpc {

constexpr int POOL_SLOTS = 4 * MAX_P;
constexpr int POOL_TABLE = 64 * (MAX_DEG + 1);  // learned table bytes kept in LDS

template <int D, bool EQ>
__device__ __forceinline__ uint2 pool_check(uint32_t *post, const uint32_t *__restrict__ ev, uint32_t *__restrict__ mp,
                                            int lp, int w, uint32_t keep, int variant, u16x2 lpe, u16x2 lpo) {
    uint32_t vi[D], pw[D], mw[D], pn[D], mn[D];
#pragma unroll
    for (int j = 0; j < D; j++) vi[j] = (ev[j] << lp) + (uint32_t)w;
#pragma unroll
    for (int j = 0; j < D; j++) pw[j] = post[vi[j]];
#pragma unroll
    for (int j = 0; j < D; j++) mw[j] = mp[j] & keep;
    const uint2 pr = check_core<D, EQ>(pw, mw, false, variant, false, lpe, lpo, pn, mn);
#pragma unroll
    for (int j = 0; j < D; j++) {
        mp[j] = mn[j];
        post[vi[j]] = pn[j];
    }
    return pr;
}

// per byte of a posterior word: 0x80 if the byte is > 0 (decision 1), else 0
__device__ __forceinline__ uint32_t positive_bytes(uint32_t x) {
    const uint32_t nz = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;  // bit 7: byte != 0
    return nz & ~x & 0x80808080u;
}

// syndrome of the decisions over every check, per byte (bit 7 of each byte
// set if some check of that frame is unsatisfied), for this thread's items
template <int UD>
__device__ __forceinline__ uint32_t syndrome_items(const PoolArgs &a, const uint32_t *post, int c_first, int cstride,
                                                   int lp, int w) {
    uint32_t synd = 0;
    if constexpr (UD > 0) {
        constexpr int H = (UD + 1) / 2;
        const char *pbase = reinterpret_cast<const char *>(post) + 4 * w;
        for (int c = c_first; c < a.C; c += cstride) {
            uint32_t off[H];
#pragma unroll
            for (int h = 0; h < H; h++) off[h] = a.edge_off16[(size_t)c * H + h];
            uint32_t par = 0;
#pragma unroll
            for (int j = 0; j < UD; j++) {
                const uint32_t o = (j & 1) ? (off[j >> 1] >> 16) : (off[j >> 1] & 0xFFFFu);
                par ^= positive_bytes(*reinterpret_cast<const uint32_t *>(pbase + o));
            }
            synd |= par;
        }
    } else {
        for (int c = c_first; c < a.C; c += cstride) {
            const uint2 ci = a.checks[c];
            const int D = (int)(ci.x & 0xFFu);
            const uint32_t *ev = a.edge_var + ci.y;
            uint32_t par = 0;
            for (int j = 0; j < D; j++) par ^= positive_bytes(post[(ev[j] << lp) + (uint32_t)w]);
            synd |= par;
        }
    }
    return synd;
}

template <bool EQ, int DMAXC>
__device__ __forceinline__ uint2 pool_check_any(int D, uint32_t *post, const uint32_t *__restrict__ ev,
                                                uint32_t *__restrict__ mp, int lp, int w, uint32_t keep, int variant,
                                                u16x2 lpe, u16x2 lpo) {
    switch (D) {
#define LDPC_PCASE(d)                                                                        \
    case d:                                                                                  \
        if constexpr (d <= DMAXC)                                                            \
            return pool_check<d, EQ>(post, ev, mp, lp, w, keep, variant, lpe, lpo);          \
        else                                                                                 \
            return make_uint2(0u, 0u);
        LDPC_PCASE(1) LDPC_PCASE(2) LDPC_PCASE(3) LDPC_PCASE(4) LDPC_PCASE(5) LDPC_PCASE(6)
        LDPC_PCASE(7) LDPC_PCASE(8) LDPC_PCASE(9) LDPC_PCASE(10) LDPC_PCASE(11) LDPC_PCASE(12)
        LDPC_PCASE(13) LDPC_PCASE(14) LDPC_PCASE(15) LDPC_PCASE(16) LDPC_PCASE(17) LDPC_PCASE(18)
        LDPC_PCASE(19) LDPC_PCASE(20) LDPC_PCASE(21) LDPC_PCASE(22) LDPC_PCASE(23) LDPC_PCASE(24)
        LDPC_PCASE(25) LDPC_PCASE(26) LDPC_PCASE(27) LDPC_PCASE(28) LDPC_PCASE(29) LDPC_PCASE(30)
        LDPC_PCASE(31) LDPC_PCASE(32)
#undef LDPC_PCASE
    default: return make_uint2(0u, 0u);
    }
}

template <bool EQ, int DMAXC, int UD, int G>
__global__ __launch_bounds__(UD > 0 ? 512 : (DMAXC <= 8 ? 1024 : 256)) void k_pool(const PoolArgs a) {
    extern __shared__ uint32_t lds_post[];
    __shared__ long long s_fid[POOL_SLOTS];
    __shared__ unsigned long long s_off[POOL_SLOTS];
    __shared__ int s_it[POOL_SLOTS];
    __shared__ int s_flag[POOL_SLOTS];  // 1: stop this iteration, 2: refilled
    __shared__ uint32_t s_keep[MAX_P];
    __shared__ uint32_t s_ove[MAX_P], s_ovo[MAX_P];
    __shared__ int8_t s_tab[POOL_TABLE];
    __shared__ int s_active;

    const int tid = threadIdx.x, T = blockDim.x, P = a.P, lp = a.log2P, N = a.N;
    const int S = 4 * P;
    const int w = tid & (P - 1);
    const int cstride = T >> lp;
    uint32_t *post = lds_post;
    uint32_t *mbase = a.msgs + (size_t)blockIdx.x * a.E * P;
    uint8_t *pb = reinterpret_cast<uint8_t *>(post);
    const bool learned = a.ltable != nullptr;
    if (learned)
        for (int i = tid; i < a.tits * a.tdeg; i += T) s_tab[i] = a.ltable[i];
    if (tid < S) {
        s_fid[tid] = -1;
        s_it[tid] = 0;
        s_flag[tid] = 1;  // every slot starts empty and asks for a frame
    }
    if (tid < P) s_keep[tid] = 0xFFFFFFFFu;
    __syncthreads();

    for (;;) {
        // ---- refill: every slot flagged 1 takes the next frame of the queue
        if (tid == 0) {
            int n = 0;
            for (int s = 0; s < S; s++) n += s_flag[s] == 1;
            const unsigned long long total = a.total_ptr ? (unsigned long long)*a.total_ptr : a.total;
            unsigned long long base = n ? atomicAdd(a.next, (unsigned long long)n) : 0ull;
            int active = 0;
            for (int s = 0; s < S; s++) {
                if (s_flag[s] == 1) {
                    if (base < total) {
                        s_fid[s] = (long long)base;  // queue position; the frame id follows below
                        s_it[s] = 0;
                        s_flag[s] = 2;
                        s_keep[s >> 2] &= ~(0xFFu << (8 * (s & 3)));
                    } else {
                        s_fid[s] = -1;
                        s_flag[s] = 0;
                    }
                    base++;
                } else {
                    s_flag[s] = 0;
                }
                active += s_fid[s] >= 0;
            }
            s_active = active;
        }
        __syncthreads();
        if (s_active == 0) break;
        // queue position -> frame id and its byte offset, one slot per thread
        if (tid < S && s_flag[tid] == 2) {
            const unsigned long long q = (unsigned long long)s_fid[tid];
            const unsigned long long f = a.frame_list ? (unsigned long long)a.frame_list[q] : q;
            s_fid[tid] = (long long)f;
            s_off[tid] = a.frame_off ? a.frame_off[f] : f * (unsigned long long)N;
        }
        __syncthreads();
        // ---- load the LLRs of the refilled frames into their lanes
        for (int s = 0; s < S; s++) {
            if (s_flag[s] != 2) continue;  // uniform: flags are read after the barrier
            const int8_t *src = a.in_frames + s_off[s];
            const int wl = s >> 2, k = s & 3;
            for (int v = tid; v < N; v += T) pb[((size_t)v * P + wl) * 4 + k] = (uint8_t)src[v];
        }
        if (tid < P) {
            s_ove[tid] = 0;
            s_ovo[tid] = 0;
        }
        __syncthreads();

        // ---- one layered iteration over all slots
        const uint32_t keep = s_keep[w];
        int rowoff[4] = {0, 0, 0, 0};
        if (learned)
#pragma unroll
            for (int k = 0; k < 4; k++) rowoff[k] = min(s_it[4 * w + k], a.tits - 1) * a.tdeg;
        uint32_t ove = 0, ovo = 0;
        if constexpr (UD > 0) {
            u16x2 lpe = splat(0u), lpo = splat(0u);
            if (learned) {
                const int col = min(UD, a.tdeg - 1);
                lpe = u2(learned_half(a.variant, s_tab[rowoff[0] + col]) |
                         (learned_half(a.variant, s_tab[rowoff[2] + col]) << 16));
                lpo = u2(learned_half(a.variant, s_tab[rowoff[1] + col]) |
                         (learned_half(a.variant, s_tab[rowoff[3] + col]) << 16));
            }
            for (int L = 0; L < a.n_levels; L++) {
                const uint2 pr = uniform_items<UD, G, EQ, false>(a.edge_off16, P, a.variant, post, mbase,
                                                                 a.level_beg[L] + (tid >> lp), a.level_beg[L + 1],
                                                                 cstride, w, keep, lpe, lpo);
                ove |= pr.x;
                ovo |= pr.y;
                __syncthreads();
            }
        } else {
            for (int L = 0; L < a.n_levels; L++) {
                const int lb = a.level_beg[L], le = a.level_beg[L + 1];
                for (int c = lb + (tid >> lp); c < le; c += cstride) {
                    const uint2 ci = a.checks[c];
                    const int D = (int)(ci.x & 0xFFu);
                    u16x2 lpe = splat(0u), lpo = splat(0u);
                    if (learned) {
                        const int col = min(D, a.tdeg - 1);
                        const uint32_t l0 = learned_half(a.variant, s_tab[rowoff[0] + col]);
                        const uint32_t l1 = learned_half(a.variant, s_tab[rowoff[1] + col]);
                        const uint32_t l2 = learned_half(a.variant, s_tab[rowoff[2] + col]);
                        const uint32_t l3 = learned_half(a.variant, s_tab[rowoff[3] + col]);
                        lpe = u2(l0 | (l2 << 16));
                        lpo = u2(l1 | (l3 << 16));
                    }
                    const uint2 pr = pool_check_any<EQ, DMAXC>(D, post, a.edge_var + ci.y,
                                                               mbase + (size_t)ci.y * P + (size_t)w * D, lp, w, keep,
                                                               a.variant, lpe, lpo);
                    ove |= pr.x;
                    ovo |= pr.y;
                }
                __syncthreads();
            }
        }
        if (a.stop_rule == 0) {  // syndrome of the decisions after the iteration
            ove = syndrome_items<UD>(a, post, tid >> lp, cstride, lp, w);
            ovo = 0;
        }
        if (ove) atomicOr(&s_ove[w], ove);
        if (ovo) atomicOr(&s_ovo[w], ovo);
        __syncthreads();

        // ---- per-frame stop decision
        if (tid < S && s_fid[tid] >= 0) {
            const int wl = tid >> 2, k = tid & 3;
            bool odd;
            if (a.stop_rule == 0) {
                odd = (s_ove[wl] >> (8 * k + 7)) & 1u;
            } else {
                const uint32_t src = (k & 1) ? s_ovo[wl] : s_ove[wl];
                odd = (src >> ((k & 2) ? 31 : 15)) & 1u;
            }
            const int it = ++s_it[tid];
            s_flag[tid] = (!odd || it >= a.max_iters) ? 1 : 0;
        }
        if (tid < P) s_keep[tid] = 0xFFFFFFFFu;
        __syncthreads();

        // ---- write the stopped frames: decisions (post > 0) and iteration count
        for (int s = 0; s < S; s++) {
            if (s_flag[s] != 1 || s_fid[s] < 0) continue;
            const size_t f = (size_t)s_fid[s];
            const int wl = s >> 2, k = s & 3;
            uint8_t *dst = a.out_hard + s_off[s];
            if ((N & 3) == 0 && (s_off[s] & 3) == 0) {
                uint32_t *d32 = reinterpret_cast<uint32_t *>(dst);
                for (int q = tid; q < (N >> 2); q += T) {
                    uint32_t x = 0;
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        x |= (uint32_t)((int8_t)pb[((size_t)(4 * q + r) * P + wl) * 4 + k] > 0) << (8 * r);
                    d32[q] = x;
                }
            } else {
                for (int v = tid; v < N; v += T) dst[v] = (uint8_t)((int8_t)pb[((size_t)v * P + wl) * 4 + k] > 0);
            }
            if (tid == 0 && a.out_iters) a.out_iters[f] = s_it[s];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Pair mode: for frames whose 4-frame posterior block does not fit in LDS
// (N * 4 bytes > the LDS budget, e.g. N = 64800), the posterior holds 2
// frames per variable in 16 bits -- exactly one packed-16 half, so the check
// update costs the same per frame-edge -- and the messages are 16-bit pairs
// too ([workgroup][level-order check][pair][j]).  A u16 (b0 | b1 << 8) becomes
// the half's register {b0 << 8, b1 << 8} with one v_perm_b32 and goes back
// with another.
// ---------------------------------------------------------------------------
__device__ __forceinline__ u16x2 pair_to_half(uint32_t v16) { return u2(__builtin_amdgcn_perm(0u, v16, 0x010C000Cu)); }
__device__ __forceinline__ uint16_t half_to_pair(u16x2 h) {
    return (uint16_t)__builtin_amdgcn_perm(0u, w32(h), 0x0C0C0301u);
}

template <int D, bool EQ>
__device__ __forceinline__ uint32_t pair_check(uint16_t *post, const uint32_t *__restrict__ ev,
                                               uint16_t *__restrict__ mp, int lp, int w, u16x2 keep, int variant,
                                               u16x2 lpar) {
    uint32_t vi[D];
    u16x2 x[D], m[D], pn[D], mn[D];
#pragma unroll
    for (int j = 0; j < D; j++) vi[j] = (ev[j] << lp) + (uint32_t)w;
#pragma unroll
    for (int j = 0; j < D; j++) x[j] = pair_to_half(post[vi[j]]);
#pragma unroll
    for (int j = 0; j < D; j++) m[j] = pair_to_half(mp[j]) & keep;
    const uint32_t par = check_half<D, EQ>(x, m, false, variant, false, lpar, pn, mn);
#pragma unroll
    for (int j = 0; j < D; j++) {
        mp[j] = half_to_pair(mn[j]);
        post[vi[j]] = half_to_pair(pn[j]);
    }
    return par;
}

template <bool EQ, int DMAXC>
__device__ __forceinline__ uint32_t pair_check_any(int D, uint16_t *post, const uint32_t *__restrict__ ev,
                                                   uint16_t *__restrict__ mp, int lp, int w, u16x2 keep, int variant,
                                                   u16x2 lpar) {
    switch (D) {
#define LDPC_QCASE(d)                                                                        \
    case d:                                                                                  \
        if constexpr (d <= DMAXC)                                                            \
            return pair_check<d, EQ>(post, ev, mp, lp, w, keep, variant, lpar);             \
        else                                                                                 \
            return 0u;
        LDPC_QCASE(1) LDPC_QCASE(2) LDPC_QCASE(3) LDPC_QCASE(4) LDPC_QCASE(5) LDPC_QCASE(6)
        LDPC_QCASE(7) LDPC_QCASE(8) LDPC_QCASE(9) LDPC_QCASE(10) LDPC_QCASE(11) LDPC_QCASE(12)
        LDPC_QCASE(13) LDPC_QCASE(14) LDPC_QCASE(15) LDPC_QCASE(16) LDPC_QCASE(17) LDPC_QCASE(18)
        LDPC_QCASE(19) LDPC_QCASE(20) LDPC_QCASE(21) LDPC_QCASE(22) LDPC_QCASE(23) LDPC_QCASE(24)
        LDPC_QCASE(25) LDPC_QCASE(26) LDPC_QCASE(27) LDPC_QCASE(28) LDPC_QCASE(29) LDPC_QCASE(30)
        LDPC_QCASE(31) LDPC_QCASE(32)
#undef LDPC_QCASE
    default: return 0u;
    }
}

template <bool EQ, int DMAXC>
__global__ __launch_bounds__(DMAXC <= 8 ? 1024 : 256) void k_pool_pairs(const PoolArgs a) {
    extern __shared__ uint16_t lds_post16[];
    __shared__ long long s_fid[POOL_SLOTS];
    __shared__ unsigned long long s_off[POOL_SLOTS];
    __shared__ int s_it[POOL_SLOTS];
    __shared__ int s_flag[POOL_SLOTS];  // 1: stop this iteration, 2: refilled
    __shared__ uint32_t s_keep[MAX_P];  // per pair: 0xFF per byte whose messages are kept
    __shared__ uint32_t s_ov[MAX_P];
    __shared__ int8_t s_tab[POOL_TABLE];
    __shared__ int s_active;

    const int tid = threadIdx.x, T = blockDim.x, P = a.P, lp = a.log2P, N = a.N;
    const int S = 2 * P;
    const int w = tid & (P - 1);
    const int cstride = T >> lp;
    uint16_t *post = lds_post16;
    uint16_t *mbase = reinterpret_cast<uint16_t *>(a.msgs) + (size_t)blockIdx.x * a.E * P;
    uint8_t *pb = reinterpret_cast<uint8_t *>(post);
    const bool learned = a.ltable != nullptr;
    if (learned)
        for (int i = tid; i < a.tits * a.tdeg; i += T) s_tab[i] = a.ltable[i];
    if (tid < S) {
        s_fid[tid] = -1;
        s_it[tid] = 0;
        s_flag[tid] = 1;
    }
    if (tid < P) s_keep[tid] = 0xFFFFu;
    __syncthreads();

    for (;;) {
        // ---- refill (as k_pool)
        if (tid == 0) {
            int n = 0;
            for (int s = 0; s < S; s++) n += s_flag[s] == 1;
            const unsigned long long total = a.total_ptr ? (unsigned long long)*a.total_ptr : a.total;
            unsigned long long base = n ? atomicAdd(a.next, (unsigned long long)n) : 0ull;
            int active = 0;
            for (int s = 0; s < S; s++) {
                if (s_flag[s] == 1) {
                    if (base < total) {
                        s_fid[s] = (long long)base;
                        s_it[s] = 0;
                        s_flag[s] = 2;
                        s_keep[s >> 1] &= ~(0xFFu << (8 * (s & 1)));
                    } else {
                        s_fid[s] = -1;
                        s_flag[s] = 0;
                    }
                    base++;
                } else {
                    s_flag[s] = 0;
                }
                active += s_fid[s] >= 0;
            }
            s_active = active;
        }
        __syncthreads();
        if (s_active == 0) break;
        if (tid < S && s_flag[tid] == 2) {
            const unsigned long long q = (unsigned long long)s_fid[tid];
            const unsigned long long f = a.frame_list ? (unsigned long long)a.frame_list[q] : q;
            s_fid[tid] = (long long)f;
            s_off[tid] = a.frame_off ? a.frame_off[f] : f * (unsigned long long)N;
        }
        __syncthreads();
        for (int s = 0; s < S; s++) {
            if (s_flag[s] != 2) continue;
            const int8_t *src = a.in_frames + s_off[s];
            const int wl = s >> 1, k = s & 1;
            for (int v = tid; v < N; v += T) pb[((size_t)v * P + wl) * 2 + k] = (uint8_t)src[v];
        }
        if (tid < P) s_ov[tid] = 0;
        __syncthreads();

        // ---- one layered iteration over all slots
        const u16x2 keep = pair_to_half(s_keep[w]);
        int rowoff[2] = {0, 0};
        if (learned)
#pragma unroll
            for (int k = 0; k < 2; k++) rowoff[k] = min(s_it[2 * w + k], a.tits - 1) * a.tdeg;
        uint32_t ov = 0;
        for (int L = 0; L < a.n_levels; L++) {
            const int lb = a.level_beg[L], le = a.level_beg[L + 1];
            for (int c = lb + (tid >> lp); c < le; c += cstride) {
                const uint2 ci = a.checks[c];
                const int D = (int)(ci.x & 0xFFu);
                u16x2 lpar = splat(0u);
                if (learned) {
                    const int col = min(D, a.tdeg - 1);
                    lpar = u2(learned_half(a.variant, s_tab[rowoff[0] + col]) |
                              (learned_half(a.variant, s_tab[rowoff[1] + col]) << 16));
                }
                ov |= pair_check_any<EQ, DMAXC>(D, post, a.edge_var + ci.y, mbase + (size_t)ci.y * P + (size_t)w * D,
                                                lp, w, keep, a.variant, lpar);
            }
            __syncthreads();
        }
        if (a.stop_rule == 0) {  // syndrome of the decisions after the iteration
            ov = 0;
            for (int c = tid >> lp; c < a.C; c += cstride) {
                const uint2 ci = a.checks[c];
                const int D = (int)(ci.x & 0xFFu);
                const uint32_t *ev = a.edge_var + ci.y;
                uint32_t par = 0;
                for (int j = 0; j < D; j++) par ^= positive_bytes(post[(ev[j] << lp) + (uint32_t)w]);
                ov |= par;
            }
        }
        if (ov) atomicOr(&s_ov[w], ov);
        __syncthreads();

        // ---- per-frame stop decision
        if (tid < S && s_fid[tid] >= 0) {
            const int wl = tid >> 1, k = tid & 1;
            const uint32_t bit = a.stop_rule == 0 ? (8u * k + 7u) : (k ? 31u : 15u);
            const bool odd = (s_ov[wl] >> bit) & 1u;
            const int it = ++s_it[tid];
            s_flag[tid] = (!odd || it >= a.max_iters) ? 1 : 0;
        }
        if (tid < P) s_keep[tid] = 0xFFFFu;
        __syncthreads();

        // ---- write the stopped frames: decisions (post > 0) and iteration count
        for (int s = 0; s < S; s++) {
            if (s_flag[s] != 1 || s_fid[s] < 0) continue;
            const size_t f = (size_t)s_fid[s];
            const int wl = s >> 1, k = s & 1;
            uint8_t *dst = a.out_hard + s_off[s];
            for (int v = tid; v < N; v += T) dst[v] = (uint8_t)((int8_t)pb[((size_t)v * P + wl) * 2 + k] > 0);
            if (tid == 0 && a.out_iters) a.out_iters[f] = s_it[s];
        }
        __syncthreads();
    }
}

template <bool EQ, int DMAXC>
static hipError_t launch_pairs_t(const PoolArgs &a, int grid, int threads, hipStream_t s) {
    hipLaunchKernelGGL((k_pool_pairs<EQ, DMAXC>), dim3(grid), dim3(threads), (size_t)a.N * a.P * 2, s, a);
    return hipGetLastError();
}

template <bool EQ>
static const void *pairs_kernel(int dmax) {
    if (dmax <= 8) return (const void *)k_pool_pairs<EQ, 8>;
    if (dmax <= 16) return (const void *)k_pool_pairs<EQ, 16>;
    return (const void *)k_pool_pairs<EQ, 32>;
}

template <bool EQ, int DMAXC, int UD, int G>
static hipError_t launch_pool_t(const PoolArgs &a, int grid, int threads, hipStream_t s) {
    hipLaunchKernelGGL((k_pool<EQ, DMAXC, UD, G>), dim3(grid), dim3(threads), (size_t)a.N * a.P * 4, s, a);
    return hipGetLastError();
}

// the instantiation that runs a pool: generic degree classes, or the
// single-degree path (ud 6..8, G items in flight per thread)
template <bool EQ>
static const void *pool_kernel(int dmax, int ud, int G) {
    if (ud == 6) return G >= 4 ? (const void *)k_pool<EQ, 8, 6, 4> : G == 2 ? (const void *)k_pool<EQ, 8, 6, 2> : (const void *)k_pool<EQ, 8, 6, 1>;
    if (ud == 7) return G >= 4 ? (const void *)k_pool<EQ, 8, 7, 4> : G == 2 ? (const void *)k_pool<EQ, 8, 7, 2> : (const void *)k_pool<EQ, 8, 7, 1>;
    if (ud == 8) return G >= 4 ? (const void *)k_pool<EQ, 8, 8, 4> : G == 2 ? (const void *)k_pool<EQ, 8, 8, 2> : (const void *)k_pool<EQ, 8, 8, 1>;
    if (dmax <= 8) return (const void *)k_pool<EQ, 8, 0, 1>;
    if (dmax <= 16) return (const void *)k_pool<EQ, 16, 0, 1>;
    return (const void *)k_pool<EQ, 32, 0, 1>;
}

template <bool EQ>
static hipError_t launch_pool_eq(const PoolArgs &a, int dmax, int ud, int G, int grid, int threads, hipStream_t s) {
#define LDPC_PL(U, GG)                                                                       \
    if (ud == U && (GG == 4 ? G >= 4 : G == GG)) return launch_pool_t<EQ, 8, U, GG>(a, grid, threads, s);
    LDPC_PL(6, 1) LDPC_PL(6, 2) LDPC_PL(6, 4) LDPC_PL(7, 1) LDPC_PL(7, 2) LDPC_PL(7, 4)
    LDPC_PL(8, 1) LDPC_PL(8, 2) LDPC_PL(8, 4)
#undef LDPC_PL
    if (ud > 0) return hipErrorInvalidValue;
    if (dmax <= 8) return launch_pool_t<EQ, 8, 0, 1>(a, grid, threads, s);
    if (dmax <= 16) return launch_pool_t<EQ, 16, 0, 1>(a, grid, threads, s);
    return launch_pool_t<EQ, 32, 0, 1>(a, grid, threads, s);
}

hipError_t launch_pool(const PoolArgs &a, int dmax, int grid, int threads, hipStream_t s) {
    if (a.pairs) {
        if (dmax < 1 || dmax > MAX_DEG || threads > max_threads_for_degree(dmax, false) || grid < 1 || a.uniform_deg)
            return hipErrorInvalidValue;
        const bool eq = a.variant == VAR_2NMS;
        if (dmax <= 8) return eq ? launch_pairs_t<true, 8>(a, grid, threads, s) : launch_pairs_t<false, 8>(a, grid, threads, s);
        if (dmax <= 16) return eq ? launch_pairs_t<true, 16>(a, grid, threads, s) : launch_pairs_t<false, 16>(a, grid, threads, s);
        return eq ? launch_pairs_t<true, 32>(a, grid, threads, s) : launch_pairs_t<false, 32>(a, grid, threads, s);
    }
    const int ud = a.uniform_deg;
    if (ud > 0 && !a.edge_off16) return hipErrorInvalidValue;
    if (dmax < 1 || dmax > MAX_DEG || threads > max_threads_for_degree(dmax, ud > 0) || grid < 1)
        return hipErrorInvalidValue;
    if (ud > 0 && (ud != dmax || !uniform_degree_supported(ud))) return hipErrorInvalidValue;
    return a.variant == VAR_2NMS ? launch_pool_eq<true>(a, dmax, ud, a.batch, grid, threads, s)
                                 : launch_pool_eq<false>(a, dmax, ud, a.batch, grid, threads, s);
}

int pool_prepare(const PoolArgs &a, int dmax, int threads) {
    int n = 0;
    const int ud = a.uniform_deg;
    const bool eq = a.variant == VAR_2NMS;
    const void *f = a.pairs ? (eq ? pairs_kernel<true>(dmax) : pairs_kernel<false>(dmax))
                            : (eq ? pool_kernel<true>(dmax, ud, a.batch) : pool_kernel<false>(dmax, ud, a.batch));
    const size_t lds = (size_t)a.N * a.P * (a.pairs ? 2 : 4);
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)POOL_LDS_DYNAMIC) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, threads, lds) != hipSuccess) return 0;
    return n;
}

__global__ void k_zero_u32(unsigned *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

hipError_t launch_zero_u32(unsigned *p, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace ldpc
