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
// A fresh lane reads zero messages in its first iteration: the per-word byte
// mask `keep` clears that lane of every message word it loads.  Learned
// tables are looked up per lane with the lane's own iteration index.
// The workgroup exits when the queue is empty and every slot is drained.
#include "check_core.h"

namespace ldpc {

// LDS by absolute byte address.  The per-edge tables of the generic paths
// (PoolArgs::edge_lds) and the pair-mode rows hold LDS addresses of the
// posterior block, which starts the kernel's dynamic LDS at address 0: the
// frame-queue kernels have no static LDS (pool_prepare checks it before a
// launch), so an access adds no base and, when P = 1, no offset at all.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) uint16_t lds_u16;
__device__ __forceinline__ lds_u32 &word_at(uint32_t addr) { return *(lds_u32 *)(uintptr_t)addr; }
__device__ __forceinline__ lds_u16 &pair_at(uint32_t addr) { return *(lds_u16 *)(uintptr_t)addr; }

// one check of degree D on a word block: ev are the LDS addresses of its
// variables' first word, w4 the thread's word offset (4w)
template <int D, bool EQ>
__device__ __forceinline__ uint2 pool_check(const uint32_t *__restrict__ ev, uint32_t *__restrict__ mp, uint32_t w4,
                                            uint32_t keep, int variant, u16x2 lpe, u16x2 lpo) {
    uint32_t vi[D], pw[D], mw[D], pn[D], mn[D];
#pragma unroll
    for (int j = 0; j < D; j++) vi[j] = ev[j] + w4;
#pragma unroll
    for (int j = 0; j < D; j++) pw[j] = word_at(vi[j]);
#pragma unroll
    for (int j = 0; j < D; j++) mw[j] = mp[j];
    const uint2 pr = check_core<D, EQ>(pw, mw, false, variant, false, lpe, lpo, pn, mn, msg_unpack(keep));
#pragma unroll
    for (int j = 0; j < D; j++) {
        mp[j] = mn[j];
        word_at(vi[j]) = pn[j];
    }
    return pr;
}

// Decision parity accumulation: acc ^ (per byte of the posterior word x, bit
// 7 set if the byte is > 0, i.e. decision 1).  The low 7 bits plus 0x7F carry
// into bit 7 exactly when they are nonzero (no carry leaves the byte), and
// & ~x clears negative bytes; the AND-NOT and the XOR into the accumulator
// are one v_bitop3 (truth table 0x9A: (u & ~x) ^ acc, table index
// u << 2 | x << 1 | acc).  Bits 0..6 of each byte are junk: keep 0x80808080
// once per check.
__device__ __forceinline__ uint32_t xor_positive(uint32_t acc, uint32_t x) {
    const uint32_t u = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
    return __builtin_amdgcn_bitop3_b32(u, x, acc, 0x9A);
}
// the same on two words at once: no carry crosses a byte, so one 64-bit add
// (v_lshl_add_u64) does the work of two 32-bit adds
__device__ __forceinline__ uint64_t xor_positive64(uint64_t acc, uint64_t x) {
    const uint64_t u = (x & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full;
    const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)u, (uint32_t)x, (uint32_t)acc, 0x9A);
    const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(u >> 32), (uint32_t)(x >> 32), (uint32_t)(acc >> 32), 0x9A);
    return ((uint64_t)hi << 32) | lo;
}

// syndrome of the decisions over every check, per byte (bit 7 of each byte
// set if some check of that frame is unsatisfied), for this thread's items.
// Single-degree path with P = 2: a thread takes a whole check and reads both
// words of each variable with one 64-bit LDS load (ov.x: word 0, ov.y: word 1).
template <int UD>
__device__ __forceinline__ uint2 syndrome_items(const PoolArgs &a, const uint32_t *post, int lp, int w) {
    const int tid = threadIdx.x, T = blockDim.x;
    static_assert(UD > 0, "the generic path tests its checks in the kernel (early exit)");
    uint2 synd = make_uint2(0u, 0u);
    {
        constexpr int H = (UD + 1) / 2;
        if (a.P == 2) {
            // SB checks per thread at a time: their offsets are loaded together
            constexpr int SB = 4;
            const char *pbase = reinterpret_cast<const char *>(post);
            const char *obase = reinterpret_cast<const char *>(a.edge_off16);
            for (int c0 = tid; c0 < a.C; c0 += SB * T) {
                uint32_t off[SB][H];
#pragma unroll
                for (int b = 0; b < SB; b++) {
                    const int c = min(c0 + b * T, a.C - 1);
                    const uint32_t *op = reinterpret_cast<const uint32_t *>(obase + __umul24(c, 4 * H));
#pragma unroll
                    for (int h = 0; h < H; h++) off[b][h] = op[h];
                }
#pragma unroll
                for (int b = 0; b < SB; b++) {
                    uint64_t pp = 0;
#pragma unroll
                    for (int j = 0; j < UD; j++) {
                        const uint32_t o = (j & 1) ? (off[b][j >> 1] >> 16) : (off[b][j >> 1] & 0xFFFFu);
                        pp = xor_positive64(pp, *reinterpret_cast<const uint64_t *>(pbase + o));
                    }
                    // a clamped index repeats check C-1: OR-ing its parity twice is harmless
                    synd.x |= (uint32_t)pp;
                    synd.y |= (uint32_t)(pp >> 32);
                    __builtin_amdgcn_sched_barrier(0);  // one check at a time (registers)
                }
            }
        } else {
            const char *pbase = reinterpret_cast<const char *>(post) + 4 * w;
            for (int c = tid >> lp; c < a.C; c += T >> lp) {
                uint32_t off[H];
#pragma unroll
                for (int h = 0; h < H; h++) off[h] = a.edge_off16[(size_t)c * H + h];
                uint32_t par = 0;
#pragma unroll
                for (int j = 0; j < UD; j++) {
                    const uint32_t o = (j & 1) ? (off[j >> 1] >> 16) : (off[j >> 1] & 0xFFFFu);
                    par = xor_positive(par, *reinterpret_cast<const uint32_t *>(pbase + o));
                }
                synd.x |= par;
            }
        }
    }
    synd.x &= 0x80808080u;
    synd.y &= 0x80808080u;
    return synd;
}

template <bool EQ, int DMAXC>
__device__ __forceinline__ uint2 pool_check_any(int D, const uint32_t *__restrict__ ev, uint32_t *__restrict__ mp,
                                                uint32_t w4, uint32_t keep, int variant, u16x2 lpe, u16x2 lpo) {
    switch (D) {
#define LDPC_PCASE(d)                                                                        \
    case d:                                                                                  \
        if constexpr (d <= DMAXC)                                                            \
            return pool_check<d, EQ>(ev, mp, w4, keep, variant, lpe, lpo);                   \
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

// Entries (degree, first edge) and variable LDS addresses of NG checks c0,
// c0 + cs, ... below `le` (a missing check gets D = 0), all loads issued
// before any is used.  DMAXC addresses are read whatever D is (those past D
// belong to the next check and are never used); the table is padded so that
// the last check's reads stay inside it.
template <int NG, int DMAXC>
__device__ __forceinline__ void gather_ids(const PoolArgs &a, int c0, int cs, int le, uint2 (&ci)[NG],
                                           uint32_t (&ev)[NG][DMAXC]) {
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int c = c0 + g * cs;
        ci[g] = c < le ? a.checks[c] : make_uint2(0u, 0u);
    }
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const uint32_t *e = a.edge_lds + ci[g].y;
#pragma unroll
        for (int j = 0; j < DMAXC; j++) ev[g][j] = e[j];
    }
}

// Slot bookkeeping of one workgroup, kept by wave 0 between barriers.
// Per round: wave 0 decides which resident frames stop (their stop test or
// the iteration limit), writes their iteration counts, and hands the free
// slots the next frames of the queue (one atomicAdd); then every thread
// writes the stopped frames' decisions and loads the new frames' LLRs for
// its own variables (the same LDS bytes, read before they are overwritten,
// so no barrier sits between the two); then the iteration runs.
struct SlotState {
    long long *fid;             // frame in the slot, -1 = empty
    unsigned long long *off;    // its byte offset
    unsigned long long *woff;   // byte offset of the frame to write out
    int *it;                    // iterations it has run
    int *io;                    // bit 0: write out, bit 1: load
    int *active;                // slots holding a frame
    uint32_t *keep, *ove, *ovo; // per word: message keep mask, stop-test bits
    int8_t *tab;                // learned table copy
};

// Dynamic LDS of a frame-queue workgroup: the posterior block, then the slot
// bookkeeping and the learned table, sized for this launch (S slots, P words)
// so that small blocks leave room for one more workgroup per CU.
// Pair mode adds one variable past the last (index N, its pairs held at zero,
// decision 0): the unused entries of a check's id row name it, so the
// syndrome scan reads every row whole, with no per-entry degree test.
__host__ __device__ inline size_t pool_post_bytes(const PoolArgs &a) {
    return ((size_t)(a.N + (a.pairs ? 1 : 0)) * a.P * (a.pairs ? 2 : 4) + 7) & ~(size_t)7;
}
__host__ __device__ inline int pool_slots(const PoolArgs &a) { return a.P * (a.pairs ? 2 : 4); }
size_t pool_lds_bytes(const PoolArgs &a) {
    const size_t S = (size_t)pool_slots(a);
    return pool_post_bytes(a) + S * 24 + S * 8 + 8 + (size_t)a.P * 12 + (a.ltable ? (size_t)a.tits * a.tdeg : 0);
}
__device__ __forceinline__ SlotState carve_slots(const PoolArgs &a, char *lds) {
    const int S = pool_slots(a);
    char *q = lds + pool_post_bytes(a);
    SlotState st;
    st.fid = reinterpret_cast<long long *>(q);
    q += 8 * S;
    st.off = reinterpret_cast<unsigned long long *>(q);
    q += 8 * S;
    st.woff = reinterpret_cast<unsigned long long *>(q);
    q += 8 * S;
    st.it = reinterpret_cast<int *>(q);
    q += 4 * S;
    st.io = reinterpret_cast<int *>(q);
    q += 4 * S;
    st.active = reinterpret_cast<int *>(q);
    q += 8;
    st.keep = reinterpret_cast<uint32_t *>(q);
    q += 4 * a.P;
    st.ove = reinterpret_cast<uint32_t *>(q);
    q += 4 * a.P;
    st.ovo = reinterpret_cast<uint32_t *>(q);
    q += 4 * a.P;
    st.tab = reinterpret_cast<int8_t *>(q);
    return st;
}

// wave 0, lane s = slot s (S <= 64): stop decisions for resident frames
// (odd_bit(s) = its stop test failed), iteration counts, and the refill from
// the queue in slot order with one atomic per round; then the per-word keep
// masks (B byte lanes per word: a refilled lane's messages are cleared) and
// the reset of the per-word stop-test bits (`reset(q)`)
static_assert(4 * MAX_P <= 64, "pool_round maps one slot to one lane of wave 0");
constexpr int STOP_NONE = 2;  // LDPC_STOP_NONE: every frame runs max_iters
template <int B, typename OddBit, typename Reset>
__device__ __forceinline__ void pool_round(const PoolArgs &a, const SlotState &st, int S, int P, OddBit odd_bit,
                                           Reset reset) {
    // opaque lane index: keeps per-lane slot addresses from being hoisted
    // out of the decoder loop (they would stay live in VGPRs through it)
    int s = threadIdx.x;
    __asm__ volatile("" : "+v"(s));
    const bool mine = s < S;
    long long fid = mine ? st.fid[s] : -1;
    unsigned long long off = mine ? st.off[s] : 0ull;
    int it = mine ? st.it[s] : 0, io = 0;
    if (fid >= 0) {
        it++;
        if (!odd_bit(s) || it >= a.max_iters) {
            if (a.out_iters) a.out_iters[fid] = it;
            io = 1;
            st.woff[s] = off;
            fid = -1;
        }
    }
    const unsigned long long free_mask = __ballot(mine && fid < 0);
    const int n_free = __popcll(free_mask);
    const unsigned long long total = a.total_ptr ? (unsigned long long)*a.total_ptr : a.total;
    // one atomic per round (a read of the head before it would add a second
    // round trip to this serial path; the head may pass `total`, harmlessly)
    unsigned long long base = total;
    if (s == 0 && n_free) base = atomicAdd(a.next, (unsigned long long)n_free);
    base = __shfl(base, 0);
    if (mine && fid < 0) {
        const unsigned long long q = base + (unsigned long long)__popcll(free_mask & ((1ull << s) - 1ull));
        if (q < total) {
            const unsigned long long f = a.frame_list ? (unsigned long long)a.frame_list[q] : q;
            fid = (long long)f;
            off = a.frame_off ? a.frame_off[f] : f * (unsigned long long)a.N;
            it = 0;
            io |= 2;
        }
    }
    if (mine) {
        st.fid[s] = fid;
        st.off[s] = off;
        st.it[s] = it;
        st.io[s] = io;
    }
    const unsigned long long active = __ballot(mine && fid >= 0);
    const unsigned long long fresh = __ballot((io & 2) != 0);
    if (s == 0) *st.active = __popcll(active);
    if (s < P) {
        uint32_t keep = B == 4 ? 0xFFFFFFFFu : 0xFFFFu;
#pragma unroll
        for (int k = 0; k < B; k++)
            if ((fresh >> (B * s + k)) & 1ull) keep &= ~(0xFFu << (8 * k));
        st.keep[s] = keep;
        reset(s);
    }
}

// every thread: decisions of the frames leaving their slots, LLRs of the
// frames entering them (byte lane k of element wl of each variable; B = bytes
// per element: 4 for words, 2 for pairs).  A thread takes dword q of a frame
// (variables 4q..4q+3).  Per slot and thread, IO_BATCH LLR dwords are loaded
// first; then per dword the leaving frame's decisions (the 4 elements read
// whole, byte k of each gathered into one dword by two v_perm and an OR, the
// positive-byte test giving the 4 decisions at once) and the entering
// frame's LLR bytes written into the same lanes.
// The variable order within a dword is rotated by (lane / 8) % 4, so that
// the lanes of a 32-lane LDS access fall on distinct banks; the gathered and
// loaded dwords are rotated to match.
constexpr int IO_BATCH = 8;
template <int B>
__device__ __forceinline__ void pool_io(const PoolArgs &a, const SlotState &st, uint8_t *pb, int S, int P) {
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));  // as in pool_round: no per-thread addresses hoisted
    const int T = blockDim.x, N = a.N;
    const uint32_t stride = (uint32_t)P * B;  // bytes per variable
    const int rot = (tid >> 3) & 3;
    for (int s = 0; s < S; s++) {
        const int io = st.io[s];
        if (!io) continue;  // uniform: read after a barrier
        const int wl = s / B, k = s % B;
        const unsigned long long wo = st.woff[s], lo = st.off[s];
        if ((N & 3) == 0 && ((wo | lo) & 3) == 0) {
            uint32_t *dst = reinterpret_cast<uint32_t *>(a.out_hard + wo);
            const uint32_t *src = reinterpret_cast<const uint32_t *>(a.in_frames + lo);
            const int Q = N >> 2;
            // element b is variable 4q + ((b + rot) & 3): its byte offset in
            // the group of 4 variables, and the gather selectors that put its
            // byte k at byte (b + rot) & 3 of the decision dword (v_perm
            // selector bytes: 0-3 pick src1, 4-7 src0, 0x0C gives zero)
            uint32_t eoff[4], sel01 = 0x0C0C0C0Cu, sel23 = 0x0C0C0C0Cu;
#pragma unroll
            for (int b = 0; b < 4; b++) eoff[b] = (uint32_t)((b + rot) & 3) * stride;
            {
                const uint32_t p0 = (uint32_t)((0 + rot) & 3), p1 = (uint32_t)((1 + rot) & 3);
                const uint32_t p2 = (uint32_t)((2 + rot) & 3), p3 = (uint32_t)((3 + rot) & 3);
                sel01 = (sel01 & ~(0xFFu << (8 * p0))) | ((uint32_t)k << (8 * p0));
                sel01 = (sel01 & ~(0xFFu << (8 * p1))) | ((uint32_t)(4 + k) << (8 * p1));
                sel23 = (sel23 & ~(0xFFu << (8 * p2))) | ((uint32_t)k << (8 * p2));
                sel23 = (sel23 & ~(0xFFu << (8 * p3))) | ((uint32_t)(4 + k) << (8 * p3));
            }
            const uint32_t lane0 = (uint32_t)wl * B;  // element of word/pair wl within a variable
            for (int q0 = tid; q0 < Q; q0 += IO_BATCH * T) {
                uint32_t y[IO_BATCH];
                if (io & 2) {
#pragma unroll
                    for (int r = 0; r < IO_BATCH; r++) {
                        const int q = q0 + r * T;
                        y[r] = q < Q ? src[q] : 0u;
                    }
                }
#pragma unroll
                for (int r = 0; r < IO_BATCH; r++) {
                    int q = q0 + r * T;
                    __asm__ volatile("" : "+v"(q));
                    if (q < Q) {
                        // the 4 elements' addresses serve the decision reads and
                        // the LLR writes (a lane's reads come before its writes)
                        uint8_t *g[4];
#pragma unroll
                        for (int b = 0; b < 4; b++) g[b] = pb + (__umul24(4 * q, stride) + lane0 + eoff[b]);
                        if (io & 1) {
                            uint32_t e[4];
#pragma unroll
                            for (int b = 0; b < 4; b++) {
                                if constexpr (B == 4)
                                    e[b] = *reinterpret_cast<const uint32_t *>(g[b]);
                                else
                                    e[b] = *reinterpret_cast<const uint16_t *>(g[b]);
                            }
                            // bytes of variables 4q..4q+3 in order, then per byte: 1 if > 0
                            const uint32_t v = __builtin_amdgcn_perm(e[1], e[0], sel01) |
                                               __builtin_amdgcn_perm(e[3], e[2], sel23);
                            dst[q] = ((((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) & ~v) >> 7) & 0x01010101u;
                        }
                        if (io & 2) {
                            // z byte b = LLR of variable 4q + ((b + rot) & 3)
                            const uint32_t z = __builtin_amdgcn_alignbyte(y[r], y[r], (uint32_t)rot);
#pragma unroll
                            for (int b = 0; b < 4; b++) g[b][k] = (uint8_t)(z >> (8 * b));
                        }
                    }
                }
            }
        } else {
            for (int v = tid; v < N; v += T) {
                uint8_t *lane = pb + (size_t)v * stride + (size_t)wl * B + k;
                if (io & 1) a.out_hard[wo + v] = (uint8_t)((int8_t)*lane > 0);
                if (io & 2) *lane = (uint8_t)a.in_frames[lo + v];
            }
        }
    }
}

template <bool EQ, int DMAXC, int UD, int G, int VK, bool P1 = false>
// Register budgets of the generic path: degrees <= 10 fit 128 VGPRs (4 waves
// per SIMD, 125 used), degrees <= 16 are held to 168 (3 waves instead of the
// 2 that 176 unconstrained registers give); neither spills.  P1: one word per
// workgroup (a.P == 1), the thread's word offset 0 at compile time.
__global__ __launch_bounds__(UD > 0 ? 512 : (DMAXC <= 8 ? 1024 : 256))
__attribute__((amdgpu_waves_per_eu(UD == 0 && DMAXC == 10 ? 4 : UD == 0 && DMAXC == 16 ? 3 : 1))) void k_pool(
    const PoolArgs a) {
    extern __shared__ uint32_t lds_post[];
    const SlotState st = carve_slots(a, reinterpret_cast<char *>(lds_post));
    uint32_t *const s_keep = st.keep, *const s_ove = st.ove, *const s_ovo = st.ovo;
    int8_t *const s_tab = st.tab;

    const int tid = threadIdx.x, T = blockDim.x, P = P1 ? 1 : a.P, lp = P1 ? 0 : a.log2P;
    const int S = 4 * P;
    const int w = tid & (P - 1);
    const uint32_t w4 = 4u * (uint32_t)w;  // byte offset of the thread's word within a variable
    const int cstride = T >> lp;
    uint32_t *post = lds_post;
    uint32_t *mbase = a.msgs + (size_t)blockIdx.x * a.E * P;
    uint8_t *pb = reinterpret_cast<uint8_t *>(post);
    const bool learned = a.ltable != nullptr;
    if (learned)
        for (int i = tid; i < a.tits * a.tdeg; i += T) s_tab[i] = a.ltable[i];
    if (tid < S) {
        st.fid[tid] = -1;
        st.it[tid] = 0;
    }
    if (tid < P) s_ove[tid] = s_ovo[tid] = 0;
    __syncthreads();

    for (;;) {
        // ---- wave 0: stop decisions, iteration counts, refill
        if (tid < 64)
            pool_round<4>(
                a, st, S, P,
                [&](int s) {
                    const int wl = s >> 2, k = s & 3;
                    if (a.stop_rule == STOP_NONE) return true;
                    if (a.stop_rule == 0) return (bool)((s_ove[wl] >> (8 * k + 7)) & 1u);
                    const uint32_t src = (k & 1) ? s_ovo[wl] : s_ove[wl];
                    return (bool)((src >> ((k & 2) ? 31 : 15)) & 1u);
                },
                [&](int q) { s_ove[q] = s_ovo[q] = 0; });
        __syncthreads();
        // ---- decisions out, LLRs in
        pool_io<4>(a, st, pb, S, P);
        __syncthreads();
        if (*st.active == 0) break;

        // ---- one layered iteration over all slots
        const uint32_t keep = s_keep[w];
        uint32_t ove = 0, ovo = 0;
        if constexpr (UD > 0) {
            int rowoff[4] = {0, 0, 0, 0};
            if (learned)
#pragma unroll
                for (int k = 0; k < 4; k++) rowoff[k] = min(st.it[4 * w + k], a.tits - 1) * a.tdeg;
            u16x2 lpe = splat(0u), lpo = splat(0u);
            if (learned) {
                const int col = min(UD, a.tdeg - 1);
                lpe = u2(learned_half(a.variant, s_tab[rowoff[0] + col]) |
                         (learned_half(a.variant, s_tab[rowoff[2] + col]) << 16));
                lpo = u2(learned_half(a.variant, s_tab[rowoff[1] + col]) |
                         (learned_half(a.variant, s_tab[rowoff[3] + col]) << 16));
            }
            int lb = a.level_beg[0], le = a.level_beg[1];
            for (int L = 0; L < a.n_levels; L++) {
                // the next level's end, loaded while this level runs
                const int ln = a.level_beg[min(L + 2, a.n_levels)];
                const uint2 pr = uniform_items<UD, G, EQ, false, VK>(a.edge_off16, P, a.variant, post, mbase,
                                                                 lb + (tid >> lp), le, cstride, w, keep, lpe, lpo);
                ove |= pr.x;
                ovo |= pr.y;
                __syncthreads();
                lb = le;
                le = ln;
            }
        } else {
            for (int L = 0; L < a.n_levels; L++) {
                const int lb = a.level_beg[L], le = a.level_beg[L + 1];
                for (int c = lb + (tid >> lp); c < le; c += cstride) {
                    const uint2 ci = a.checks[c];
                    const int D = (int)(ci.x & 0xFFu);
                    u16x2 lpe = splat(0u), lpo = splat(0u);
                    if (learned) {
                        // the lanes' table rows re-read here: no registers held across the level loop
                        const int col = min(D, a.tdeg - 1);
                        uint32_t l[4];
#pragma unroll
                        for (int k = 0; k < 4; k++)
                            l[k] = learned_half(a.variant, s_tab[min(st.it[4 * w + k], a.tits - 1) * a.tdeg + col]);
                        lpe = u2(l[0] | (l[2] << 16));
                        lpo = u2(l[1] | (l[3] << 16));
                    }
                    const uint2 pr = pool_check_any<EQ, DMAXC>(D, a.edge_lds + ci.y,
                                                               mbase + (size_t)ci.y * P + (size_t)w * D, w4, keep,
                                                               VK >= 0 ? VK : a.variant, lpe, lpo);
                    ove |= pr.x;
                    ovo |= pr.y;
                }
                __syncthreads();
            }
        }
        if (a.stop_rule == STOP_NONE) {
            // no stop test: the frame runs to max_iters
        } else if (a.stop_rule != 0) {
            if (ove) atomicOr(&s_ove[w], ove);
            if (ovo) atomicOr(&s_ovo[w], ovo);
        } else if constexpr (UD == 0) {
            // syndrome, generic path: as in pair mode, found failing checks
            // are published at once and a thread stops scanning when every
            // frame of its word already has one
            constexpr int SG = DMAXC <= 8 ? 4 : DMAXC <= 16 ? 2 : 1;
            uint32_t need = 0;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (st.fid[4 * w + k] >= 0) need |= 0x80u << (8 * k);
            uint32_t seen = 0;
            for (int c0 = tid >> lp; c0 < a.C; c0 += SG * cstride) {
                if ((s_ove[w] & need) == need) break;
                uint2 ci[SG];
                uint32_t ev[SG][DMAXC];
                gather_ids<SG, DMAXC>(a, c0, cstride, a.C, ci, ev);
                uint32_t found = 0;
#pragma unroll
                for (int g = 0; g < SG; g++) {
                    const int D = (int)(ci[g].x & 0xFFu);
                    uint32_t par = 0;
#pragma unroll
                    for (int j = 0; j < DMAXC; j++)
                        if (j < D) par = xor_positive(par, word_at(ev[g][j] + w4));
                    found |= par & 0x80808080u;
                    __builtin_amdgcn_sched_barrier(0);  // one check at a time (registers)
                }
                if (found & ~seen) atomicOr(&s_ove[w], found);
                seen |= found;
            }
        } else {  // syndrome of the decisions after the iteration
            const uint2 sy = syndrome_items<UD>(a, post, lp, w);
            if (P == 2) {  // per check: both words; one atomic per wave and word
                uint32_t r0 = 0, r1 = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t bit = 1u << (8 * k + 7);
                    if (__ballot((sy.x & bit) != 0)) r0 |= bit;
                    if (__ballot((sy.y & bit) != 0)) r1 |= bit;
                }
                if ((tid & 63) == 0) {
                    if (r0) atomicOr(&s_ove[0], r0);
                    if (r1) atomicOr(&s_ove[1], r1);
                }
            } else if (sy.x) {
                atomicOr(&s_ove[w], sy.x);
            }
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

// One pair-mode check whose variable LDS addresses and messages are already
// in registers (ev[j], mr[j] for j < D); mp is where its messages go back.
template <int D, bool EQ, int DM>
__device__ __forceinline__ uint32_t pair_check(const uint32_t (&ev)[DM], const uint32_t (&mr)[DM],
                                               uint16_t *__restrict__ mp, uint32_t w2, u16x2 keep, int variant,
                                               u16x2 lpar) {
    uint32_t vi[D];
    u16x2 x[D], m[D], pn[D], mn[D];
#pragma unroll
    for (int j = 0; j < D; j++) vi[j] = ev[j] + w2;
#pragma unroll
    for (int j = 0; j < D; j++) x[j] = pair_to_half(pair_at(vi[j]));
#pragma unroll
    for (int j = 0; j < D; j++) m[j] = pair_to_half(mr[j]) & keep;
    const uint32_t par = check_half<D, EQ>(x, m, false, variant, false, lpar, pn, mn);
#pragma unroll
    for (int j = 0; j < D; j++) {
        mp[j] = half_to_pair(mn[j]);
        pair_at(vi[j]) = half_to_pair(pn[j]);
    }
    return par;
}

template <bool EQ, int DMAXC>
__device__ __forceinline__ uint32_t pair_check_any(int D, const uint32_t (&ev)[DMAXC], const uint32_t (&mr)[DMAXC],
                                                   uint16_t *__restrict__ mp, uint32_t w2, u16x2 keep, int variant,
                                                   u16x2 lpar) {
    switch (D) {
#define LDPC_QCASE(d)                                                                        \
    case d:                                                                                  \
        if constexpr (d <= DMAXC)                                                            \
            return pair_check<d, EQ, DMAXC>(ev, mr, mp, w2, keep, variant, lpar);            \
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

// Variable ids and 16-bit pair messages of NG checks c0, c0 + cs, ... below
// `le`, all loads issued before any is used: one memory latency for the
// group instead of two per check.  A missing check gets D = 0.  DMAXC
// entries are read whatever D is (entries j >= D belong to the next check
// and are never used); the edge-id and message arrays are padded so that the
// last check's reads stay inside them.
template <int NG, int DMAXC>
__device__ __forceinline__ void pair_gather(const PoolArgs &a, const uint16_t *mbase, int c0, int cs, int le, int P,
                                            int w, uint2 (&ci)[NG], uint32_t (&ev)[NG][DMAXC],
                                            uint32_t (&mr)[NG][DMAXC]) {
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int c = c0 + g * cs;
        ci[g] = c < le ? a.checks[c] : make_uint2(0u, 0u);
    }
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int D = (int)(ci[g].x & 0xFFu);
        const uint32_t *e = a.edge_lds + ci[g].y;
        const uint16_t *m = mbase + (size_t)ci[g].y * P + (size_t)w * D;
#pragma unroll
        for (int j = 0; j < DMAXC; j++) {
            ev[g][j] = e[j];
            mr[g][j] = (uint32_t)m[j];
        }
    }
}

// Pair mode with degrees <= 8 (rows): a check's variables are one row of 8
// LDS byte offsets (variable * 2P: the pair of a thread with w = 0; a thread
// adds 2w) and its messages one 16-byte run of 8 u16 per pair, [workgroup]
// [check][pair][8], both addressed by the check index alone: two 16-byte
// loads for the offsets, one load and one store for the messages (entries
// past D unused).
template <int NG>
__device__ __forceinline__ void rows_gather(const PoolArgs &a, const uint16_t *mbase, int c0, int cs, int le, int P,
                                            int w, uint2 (&ci)[NG], uint32_t (&ev)[NG][8], uint32_t (&mr)[NG][8]) {
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int c = c0 + g * cs;
        const bool ok = c < le;
        const int cc = ok ? c : 0;
        ci[g] = ok ? a.checks[c] : make_uint2(0u, 0u);
        const uint4 *r = reinterpret_cast<const uint4 *>(a.rows + (size_t)cc * 8);
        const uint4 r0 = r[0], r1 = r[1];
        ev[g][0] = r0.x, ev[g][1] = r0.y, ev[g][2] = r0.z, ev[g][3] = r0.w;
        ev[g][4] = r1.x, ev[g][5] = r1.y, ev[g][6] = r1.z, ev[g][7] = r1.w;
        const uint4 m = *reinterpret_cast<const uint4 *>(mbase + ((size_t)cc * P + w) * 8);
        const uint32_t mw[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            mr[g][2 * k] = mw[k] & 0xFFFFu;
            mr[g][2 * k + 1] = mw[k] >> 16;
        }
    }
}
template <int NG>
__device__ __forceinline__ void rows_ids(const PoolArgs &a, int c0, int cs, int le, uint2 (&ci)[NG],
                                         uint32_t (&ev)[NG][8]) {
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int c = c0 + g * cs;
        const bool ok = c < le;
        ci[g] = ok ? a.checks[c] : make_uint2(0u, 0u);
        const uint4 *r = reinterpret_cast<const uint4 *>(a.rows + (size_t)(ok ? c : 0) * 8);
        const uint4 r0 = r[0], r1 = r[1];
        ev[g][0] = r0.x, ev[g][1] = r0.y, ev[g][2] = r0.z, ev[g][3] = r0.w;
        ev[g][4] = r1.x, ev[g][5] = r1.y, ev[g][6] = r1.z, ev[g][7] = r1.w;
    }
}
// one rows-layout check: as pair_check, the 8 messages stored as one run
// (two halves' 16-bit pairs per dword: bytes 1 and 3 of each half)
template <int D, bool EQ>
__device__ __forceinline__ uint32_t rows_check(const uint32_t (&ev)[8], const uint32_t (&mr)[8],
                                               uint4 *__restrict__ mp, uint32_t w2, u16x2 keep, int variant,
                                               u16x2 lpar) {
    uint32_t vi[D];
    u16x2 x[D], m[D], pn[D], mn[D];
#pragma unroll
    for (int j = 0; j < D; j++) vi[j] = ev[j] + w2;
#pragma unroll
    for (int j = 0; j < D; j++) x[j] = pair_to_half(pair_at(vi[j]));
#pragma unroll
    for (int j = 0; j < D; j++) m[j] = pair_to_half(mr[j]) & keep;
    const uint32_t par = check_half<D, EQ>(x, m, false, variant, false, lpar, pn, mn);
    uint32_t h[8];
#pragma unroll
    for (int j = 0; j < 8; j++) h[j] = j < D ? w32(mn[j]) : 0u;
    uint4 o;
    o.x = __builtin_amdgcn_perm(h[1], h[0], 0x07050301u);
    o.y = __builtin_amdgcn_perm(h[3], h[2], 0x07050301u);
    o.z = __builtin_amdgcn_perm(h[5], h[4], 0x07050301u);
    o.w = __builtin_amdgcn_perm(h[7], h[6], 0x07050301u);
    *mp = o;
#pragma unroll
    for (int j = 0; j < D; j++) pair_at(vi[j]) = half_to_pair(pn[j]);
    return par;
}
// One batch of NG rows-layout checks c0, c0 + cs, ... of degree D, all valid,
// in straight-line code: all loads, then all posterior gathers, then the
// check updates and stores (checks of one level share no variable, so their
// order is free).  sel_lo / sel_hi turn the 16-bit message pair in the low /
// high half of a loaded word into a packed-16 half with the lanes of frames
// that start now zeroed (the keep mask folded into the unpack).
template <int D, int NG, bool EQ>
__device__ __forceinline__ uint32_t rows_batch(const PoolArgs &a, uint16_t *mbase, int c0, int cs,
                                               int P, int w, uint32_t w2, uint32_t sel_lo, uint32_t sel_hi,
                                               int variant, u16x2 lpar) {
    uint32_t ev[NG][8], mw[NG][4];
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const int c = c0 + g * cs;
        const uint4 *r = reinterpret_cast<const uint4 *>(a.rows + (size_t)c * 8);
        const uint4 r0 = r[0], r1 = r[1];
        ev[g][0] = r0.x, ev[g][1] = r0.y, ev[g][2] = r0.z, ev[g][3] = r0.w;
        ev[g][4] = r1.x, ev[g][5] = r1.y, ev[g][6] = r1.z, ev[g][7] = r1.w;
        const uint4 m = *reinterpret_cast<const uint4 *>(mbase + ((size_t)c * P + w) * 8);
        mw[g][0] = m.x, mw[g][1] = m.y, mw[g][2] = m.z, mw[g][3] = m.w;
    }
    uint32_t vi[NG][D];
    u16x2 x[NG][D];
#pragma unroll
    for (int g = 0; g < NG; g++)
#pragma unroll
        for (int j = 0; j < D; j++) {
            vi[g][j] = ev[g][j] + w2;
            x[g][j] = pair_to_half(pair_at(vi[g][j]));
        }
    uint32_t ov = 0;
#pragma unroll
    for (int g = 0; g < NG; g++) {
        u16x2 m[D], pn[D], mn[D];
#pragma unroll
        for (int j = 0; j < D; j++) m[j] = u2(__builtin_amdgcn_perm(0u, mw[g][j >> 1], (j & 1) ? sel_hi : sel_lo));
        ov |= check_half<D, EQ>(x[g], m, false, variant, false, lpar, pn, mn);
        uint32_t h[8];
#pragma unroll
        for (int j = 0; j < 8; j++) h[j] = j < D ? w32(mn[j]) : 0u;
        uint4 o;
        o.x = __builtin_amdgcn_perm(h[1], h[0], 0x07050301u);
        o.y = __builtin_amdgcn_perm(h[3], h[2], 0x07050301u);
        o.z = __builtin_amdgcn_perm(h[5], h[4], 0x07050301u);
        o.w = __builtin_amdgcn_perm(h[7], h[6], 0x07050301u);
        *reinterpret_cast<uint4 *>(mbase + ((size_t)(c0 + g * cs) * P + w) * 8) = o;
#pragma unroll
        for (int j = 0; j < D; j++) pair_at(vi[g][j]) = half_to_pair(pn[j]);
    }
    return ov;
}

// One level whose checks all have degree D (the plan's level_deg): a thread
// takes its checks c0, c0 + cs, ... in batches of NG while NG of them remain,
// then one at a time, so no lane computes a check it does not own.
template <int D, int NG, bool EQ>
__device__ __forceinline__ uint32_t rows_level(const PoolArgs &a, uint16_t *mbase, int lb, int le,
                                               int cs, int P, int lp, int w, uint32_t w2, uint32_t sel_lo,
                                               uint32_t sel_hi, int variant, u16x2 lpar) {
    uint32_t ov = 0;
    int c0 = lb + ((int)threadIdx.x >> lp);
    for (; c0 + (NG - 1) * cs < le; c0 += NG * cs)
        ov |= rows_batch<D, NG, EQ>(a, mbase, c0, cs, P, w, w2, sel_lo, sel_hi, variant, lpar);
    for (; c0 < le; c0 += cs)
        ov |= rows_batch<D, 1, EQ>(a, mbase, c0, cs, P, w, w2, sel_lo, sel_hi, variant, lpar);
    return ov;
}

template <bool EQ>
__device__ __forceinline__ uint32_t rows_check_any(int D, const uint32_t (&ev)[8],
                                                   const uint32_t (&mr)[8], uint4 *__restrict__ mp, uint32_t w2,
                                                   u16x2 keep, int variant, u16x2 lpar) {
    switch (D) {
#define LDPC_RCASE(d) \
    case d: return rows_check<d, EQ>(ev, mr, mp, w2, keep, variant, lpar);
        LDPC_RCASE(1) LDPC_RCASE(2) LDPC_RCASE(3) LDPC_RCASE(4) LDPC_RCASE(5) LDPC_RCASE(6) LDPC_RCASE(7)
        LDPC_RCASE(8)
#undef LDPC_RCASE
    default: return 0u;
    }
}

// P1: one pair per workgroup (a.P == 1), so a thread's pair offset w is 0 at
// compile time and a rows-layout entry is the LDS address itself.  VK >= 0:
// the variant known at compile time (the magnitude map becomes straight-line
// code; a runtime switch splits every check into basic blocks), -1: a.variant.
template <bool EQ, int DMAXC, bool P1, int VK = -1>
__global__ __launch_bounds__(DMAXC <= 8 ? 1024 : 256) void k_pool_pairs(const PoolArgs a) {
    extern __shared__ uint16_t lds_post16[];
    const SlotState st = carve_slots(a, reinterpret_cast<char *>(lds_post16));
    uint32_t *const s_keep = st.keep;  // per pair: 0xFF per byte whose messages are kept
    uint32_t *const s_ov = st.ove;
    int8_t *const s_tab = st.tab;

    const int tid = threadIdx.x, T = blockDim.x, P = P1 ? 1 : a.P, lp = P1 ? 0 : a.log2P;
    const int S = 2 * P;
    const int variant = VK >= 0 ? VK : a.variant;
    const int w = tid & (P - 1);
    const uint32_t w2 = 2u * (uint32_t)w;  // byte offset of the thread's pair within a variable
    const int cstride = T >> lp;
    uint16_t *post = lds_post16;
    constexpr bool ROWS = DMAXC == 8;  // degrees <= 8: the rows layout (the host provides a.rows)
    uint16_t *mbase = reinterpret_cast<uint16_t *>(a.msgs) + (size_t)blockIdx.x * (ROWS ? (size_t)a.C * 8 : a.E) * P;
    uint8_t *pb = reinterpret_cast<uint8_t *>(post);
    const bool learned = a.ltable != nullptr;
    if (learned)
        for (int i = tid; i < a.tits * a.tdeg; i += T) s_tab[i] = a.ltable[i];
    if (tid < S) {
        st.fid[tid] = -1;
        st.it[tid] = 0;
    }
    if (tid < P) {
        s_ov[tid] = 0;
        post[(uint32_t)a.N * P + tid] = 0;  // the padding variable
    }
    __syncthreads();

    for (;;) {
        // ---- wave 0: stop decisions, iteration counts, refill
        if (tid < 64)
            pool_round<2>(
                a, st, S, P,
                [&](int s) {
                    if (a.stop_rule == STOP_NONE) return true;
                    const uint32_t bit = a.stop_rule == 0 ? (8u * (s & 1) + 7u) : ((s & 1) ? 31u : 15u);
                    return (bool)((s_ov[s >> 1] >> bit) & 1u);
                },
                [&](int q) { s_ov[q] = 0; });
        __syncthreads();
        // ---- decisions out, LLRs in
        pool_io<2>(a, st, pb, S, P);
        __syncthreads();
        if (*st.active == 0) break;

        // ---- one layered iteration over all slots
        const u16x2 keep = pair_to_half(s_keep[w]);
        int rowoff[2] = {0, 0};
        if (learned)
#pragma unroll
            for (int k = 0; k < 2; k++) rowoff[k] = min(st.it[2 * w + k], a.tits - 1) * a.tdeg;
        uint32_t ov = 0;
        constexpr int NG = DMAXC <= 8 ? 2 : 1;  // checks in flight per thread
        // message pair -> packed-16 half with fresh frames' lanes zeroed
        // (the pair in bytes 0-1 or 2-3 of a loaded word)
        uint32_t sel_lo = 0x010C000Cu;
        {
            const uint32_t k16 = s_keep[w];
            if (!(k16 & 0x00FFu)) sel_lo = (sel_lo & 0xFFFF00FFu) | 0x00000C00u;
            if (!(k16 & 0xFF00u)) sel_lo = (sel_lo & 0x00FFFFFFu) | 0x0C000000u;
        }
        const uint32_t sel_hi = sel_lo + (((sel_lo & 0x0000FF00u) != 0x0C00u ? 0x0200u : 0u) |
                                          ((sel_lo & 0xFF000000u) != 0x0C000000u ? 0x02000000u : 0u));
        for (int L = 0; L < a.n_levels; L++) {
            const int lb = a.level_beg[L], le = a.level_beg[L + 1];
            if constexpr (ROWS) {
                // a level of one degree: straight-line checks (no per-check dispatch)
                const int ud = a.level_deg ? a.level_deg[L] : 0;
                if (ud >= 4) {
                    u16x2 lpar = splat(0u);
                    if (learned) {
                        const int col = min(ud, a.tdeg - 1);
                        lpar = u2(learned_half(a.variant, s_tab[rowoff[0] + col]) |
                                  (learned_half(a.variant, s_tab[rowoff[1] + col]) << 16));
                    }
                    switch (ud) {
#define LDPC_RLEVEL(d)                                                                                          \
    case d:                                                                                                     \
        ov |= rows_level<d, NG, EQ>(a, mbase, lb, le, cstride, P, lp, w, w2, sel_lo, sel_hi, variant, lpar);   \
        break;
                        LDPC_RLEVEL(4) LDPC_RLEVEL(5) LDPC_RLEVEL(6) LDPC_RLEVEL(7) LDPC_RLEVEL(8)
#undef LDPC_RLEVEL
                    }
                    __syncthreads();
                    continue;
                }
            }
            for (int c0 = lb + (tid >> lp); c0 < le; c0 += NG * cstride) {
                uint2 ci[NG];
                uint32_t ev[NG][DMAXC], mr[NG][DMAXC];
                if constexpr (ROWS)
                    rows_gather<NG>(a, mbase, c0, cstride, le, P, w, ci, ev, mr);
                else
                    pair_gather<NG, DMAXC>(a, mbase, c0, cstride, le, P, w, ci, ev, mr);
#pragma unroll
                for (int g = 0; g < NG; g++) {
                    const int D = (int)(ci[g].x & 0xFFu);
                    if (D == 0) continue;
                    u16x2 lpar = splat(0u);
                    if (learned) {
                        const int col = min(D, a.tdeg - 1);
                        lpar = u2(learned_half(a.variant, s_tab[rowoff[0] + col]) |
                                  (learned_half(a.variant, s_tab[rowoff[1] + col]) << 16));
                    }
                    if constexpr (ROWS)
                        ov |= rows_check_any<EQ>(D, ev[g], mr[g],
                                                 reinterpret_cast<uint4 *>(mbase + ((size_t)(c0 + g * cstride) * P + w) * 8),
                                                 w2, keep, variant, lpar);
                    else
                        ov |= pair_check_any<EQ, DMAXC>(D, ev[g], mr[g],
                                                        mbase + (size_t)ci[g].y * P + (size_t)w * D, w2, keep,
                                                        variant, lpar);
                }
            }
            __syncthreads();
        }
        if (a.stop_rule == 0) {  // syndrome of the decisions after the iteration
            // The stop test only needs, per frame, whether SOME check fails:
            // found bits are published at once, and a thread stops scanning
            // when every frame of its pair already has a failing check (the
            // result for frames still without one is the full scan's).
            constexpr int SG = DMAXC <= 8 ? 4 : 2;
            const uint32_t need = (st.fid[2 * w] >= 0 ? 0x0080u : 0u) | (st.fid[2 * w + 1] >= 0 ? 0x8000u : 0u);
            ov = 0;
            for (int c0 = tid >> lp; c0 < a.C; c0 += SG * cstride) {
                if ((s_ov[w] & need) == need) break;
                uint2 ci[SG];
                uint32_t ev[SG][DMAXC];
                if constexpr (ROWS)
                    rows_ids<SG>(a, c0, cstride, a.C, ci, ev);
                else
                    gather_ids<SG, DMAXC>(a, c0, cstride, a.C, ci, ev);
                uint32_t found = 0;
#pragma unroll
                for (int g = 0; g < SG; g++) {
                    uint32_t par = 0;
                    if constexpr (ROWS) {
                        // two pairs per test (bytes 0-1: entry j, 2-3: entry j + 1), the
                        // halves folded once per check; entries past the degree name
                        // the padding variable (decision 0)
#pragma unroll
                        for (int j = 0; j < 8; j += 2)
                            par = xor_positive(par, (uint32_t)pair_at(ev[g][j] + w2) |
                                                        ((uint32_t)pair_at(ev[g][j + 1] + w2) << 16));
                        // (par ^ par >> 16) & 0x8080 (truth table 0x28: (a ^ b) & c)
                        found |= __builtin_amdgcn_bitop3_b32(par, par >> 16, 0x8080u, 0x28);
                    } else {
                        const int D = (int)(ci[g].x & 0xFFu);
#pragma unroll
                        for (int j = 0; j < DMAXC; j++)
                            if (j < D) par = xor_positive(par, pair_at(ev[g][j] + w2));
                        found |= par & 0x8080u;
                    }
                    __builtin_amdgcn_sched_barrier(0);  // one check at a time (registers)
                }
                if (found & ~ov) atomicOr(&s_ov[w], found);
                ov |= found;
            }
        } else if (ov && a.stop_rule != STOP_NONE) {
            atomicOr(&s_ov[w], ov);
        }
        __syncthreads();
    }
}

template <bool EQ, int VK>
static const void *pairs_rows_kernel(int P) {
    return P == 1 ? (const void *)k_pool_pairs<EQ, 8, true, VK> : (const void *)k_pool_pairs<EQ, 8, false, VK>;
}
template <bool EQ>
static const void *pairs_kernel(int dmax, int P, int variant) {
    if (dmax <= 8) {
        // the OMS family (OMS takes the clamp everywhere in a frame queue) and
        // learned offsets with their magnitude map compiled in
        if constexpr (!EQ) {
            if (variant == VAR_OMS || variant == VAR_OMS_V2) return pairs_rows_kernel<EQ, VAR_OMS>(P);
            if (variant == VAR_LOMS) return pairs_rows_kernel<EQ, VAR_LOMS>(P);
        }
        return pairs_rows_kernel<EQ, -1>(P);
    }
    if (dmax <= 16) return (const void *)k_pool_pairs<EQ, 16, false>;
    return (const void *)k_pool_pairs<EQ, 32, false>;
}

template <bool EQ, int VK>
static const void *uniform_kernel(int ud, int G) {
#define LDPC_UK(U)                                                                                    \
    if (ud == U) return G >= 4 ? (const void *)k_pool<EQ, 8, U, 4, VK>                               \
                               : G == 2 ? (const void *)k_pool<EQ, 8, U, 2, VK> : (const void *)k_pool<EQ, 8, U, 1, VK>;
    LDPC_UK(6) LDPC_UK(7) LDPC_UK(8)
#undef LDPC_UK
    return nullptr;
}

// The instantiation that runs a pool: pair mode by degree class; the
// single-degree path (degree 6..8, G items in flight per thread), with the
// plain OMS family and learned offsets compiled straight-line; or the generic
// degree classes.
static const void *pool_kernel(const PoolArgs &a, int dmax) {
    const bool eq = a.variant == VAR_2NMS;
    if (a.pairs) return eq ? pairs_kernel<true>(dmax, a.P, a.variant) : pairs_kernel<false>(dmax, a.P, a.variant);
    if (a.uniform_deg) {
        if (!uniform_degree_supported(a.uniform_deg) || a.uniform_deg != dmax) return nullptr;
        if (eq) return uniform_kernel<true, -1>(a.uniform_deg, a.batch);
        if (a.variant == VAR_OMS || a.variant == VAR_OMS_V2) return uniform_kernel<false, VAR_OMS>(a.uniform_deg, a.batch);
        if (a.variant == VAR_LOMS) return uniform_kernel<false, VAR_LOMS>(a.uniform_deg, a.batch);
        return uniform_kernel<false, -1>(a.uniform_deg, a.batch);
    }
    const bool p1 = a.P == 1;
    // generic degree classes; the OMS family (OMS takes the clamp everywhere
    // in a frame queue) with its magnitude map compiled in
    const bool oms = a.variant == VAR_OMS || a.variant == VAR_OMS_V2;
#define LDPC_GK(DM)                                                                                     \
    if (dmax <= DM)                                                                                     \
        return eq ? (p1 ? (const void *)k_pool<true, DM, 0, 1, -1, true> : (const void *)k_pool<true, DM, 0, 1, -1>) \
             : oms ? (p1 ? (const void *)k_pool<false, DM, 0, 1, VAR_OMS, true>                         \
                         : (const void *)k_pool<false, DM, 0, 1, VAR_OMS>)                              \
                   : (p1 ? (const void *)k_pool<false, DM, 0, 1, -1, true> : (const void *)k_pool<false, DM, 0, 1, -1>);
    LDPC_GK(8) LDPC_GK(10) LDPC_GK(16)
#undef LDPC_GK
    return eq ? (const void *)k_pool<true, 32, 0, 1, -1> : (const void *)k_pool<false, 32, 0, 1, -1>;
}

hipError_t launch_pool(const PoolArgs &a, int dmax, int grid, int threads, hipStream_t s) {
    if (dmax < 1 || dmax > MAX_DEG || grid < 1 || threads > max_threads_for_degree(dmax, a.uniform_deg > 0))
        return hipErrorInvalidValue;
    if (a.uniform_deg && !a.edge_off16) return hipErrorInvalidValue;  // the path reads its offset table
    if (a.pairs && dmax <= 8 && !a.rows) return hipErrorInvalidValue;  // pair mode, degrees <= 8: rows
    const void *f = pool_kernel(a, dmax);
    if (!f) return hipErrorInvalidValue;
    void *args[] = {const_cast<PoolArgs *>(&a)};
    return hipLaunchKernel(f, dim3(grid), dim3(threads), args, pool_lds_bytes(a), s);
}

int pool_prepare(const PoolArgs &a, int dmax, int threads) {
    int n = 0;
    const void *f = pool_kernel(a, dmax);
    if (!f) return 0;
    {
        // the edge tables and rows address LDS absolutely from 0 (word_at,
        // pair_at): no static LDS may sit in front of the dynamic block
        hipFuncAttributes fa;
        if (hipFuncGetAttributes(&fa, f) != hipSuccess || fa.sharedSizeBytes != 0) return -1;
    }
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)MAX_LDS_BYTES) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, threads, pool_lds_bytes(a)) != hipSuccess) return 0;
    return n;
}

__global__ void k_zero_u32(unsigned *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

hipError_t launch_zero_u32(unsigned *p, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, p, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Queue order: least reliable frames first.  A frame-queue launch ends when
// its last frames do; taken in input order, the last frames of a batch can
// need several times the mean iteration count while the other workgroups have
// drained.  The sum of squares of a frame's first ORDER_SAMPLE LLRs grows with
// its noise power, so frames are queued by it, largest first: a counting sort
// over ORDER_BUCKETS buckets spread over the batch's own key range (frames of
// one bucket in any order).  The frames that run longest start first and a
// batch ends on short ones; each frame decodes exactly as in any order.
// ---------------------------------------------------------------------------
constexpr int ORDER_SAMPLE = 4096, ORDER_BUCKETS = 1024;

// ORDER_FPW frames per wave, 4 waves per workgroup: every frame's key, and
// the batch's key range (one atomic pair per workgroup: a pair per frame on
// two addresses serialises in L2, about 24 ns a frame)
constexpr int ORDER_FPW = 4;
__global__ __launch_bounds__(256) void k_order_keys(const int8_t *__restrict__ in, int N, unsigned long long frames,
                                                    unsigned *__restrict__ key, unsigned *__restrict__ range) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = min(N, ORDER_SAMPLE);
    const bool dwords = (N & 3) == 0 && (reinterpret_cast<uintptr_t>(in) & 3) == 0;
    unsigned lo = 0xFFFFFFFFu, hi = 0u;
    for (int j = 0; j < ORDER_FPW; j++) {
        const unsigned long long f = ((unsigned long long)blockIdx.x * 4 + wave) * ORDER_FPW + j;
        if (f >= frames) break;  // wave-uniform
        const int8_t *x = in + f * (unsigned long long)N;
        unsigned acc = 0;
        if (dwords) {
            const uint32_t *x4 = reinterpret_cast<const uint32_t *>(x);
#pragma unroll 8
            for (int q = lane; q < (n >> 2); q += 64) {
                const uint32_t w = x4[q];
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const int y = (int8_t)(w >> (8 * b));
                    acc += (unsigned)(y * y);
                }
            }
        } else {
            for (int v = lane; v < n; v += 64) {
                const int y = x[v];
                acc += (unsigned)(y * y);
            }
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d);  // <= 4096 * 16384 = 2^26
        if (lane == 0) key[f] = acc;
        lo = min(lo, acc);
        hi = max(hi, acc);
    }
    __shared__ unsigned part[2][4];
    if (lane == 0) {
        part[0][wave] = lo;
        part[1][wave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        lo = min(min(part[0][0], part[0][1]), min(part[0][2], part[0][3]));
        hi = max(max(part[1][0], part[1][1]), max(part[1][2], part[1][3]));
        if (lo <= hi) {  // the workgroup had a frame
            atomicMin(&range[0], lo);
            atomicMax(&range[1], hi);
        }
    }
}

// bucket of a key: 0 for the largest key of the batch
__device__ __forceinline__ unsigned order_bucket(unsigned k, const unsigned *range) {
    const unsigned long long span = (unsigned long long)(range[1] - range[0]) + 1ull;
    return (unsigned)(ORDER_BUCKETS - 1) - (unsigned)((unsigned long long)(k - range[0]) * ORDER_BUCKETS / span);
}

__global__ void k_order_count(const unsigned *__restrict__ key, unsigned long long frames,
                              const unsigned *__restrict__ range, unsigned *__restrict__ counts) {
    const unsigned long long f = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < frames) atomicAdd(&counts[order_bucket(key[f], range)], 1u);
}

// exclusive prefix sum of the bucket counts into the bucket starts (one workgroup)
__global__ __launch_bounds__(ORDER_BUCKETS) void k_order_scan(const unsigned *__restrict__ counts,
                                                             unsigned *__restrict__ starts) {
    __shared__ unsigned c[ORDER_BUCKETS];
    const int t = threadIdx.x;
    c[t] = counts[t];
    __syncthreads();
    for (int d = 1; d < ORDER_BUCKETS; d <<= 1) {
        const unsigned v = t >= d ? c[t - d] : 0u;
        __syncthreads();
        c[t] += v;
        __syncthreads();
    }
    starts[t] = c[t] - counts[t];
}

// queue position of every frame: its bucket's start plus its rank there
__global__ void k_order_scatter(const unsigned *__restrict__ key, unsigned long long frames,
                                const unsigned *__restrict__ range, unsigned *__restrict__ starts,
                                unsigned *__restrict__ list) {
    const unsigned long long f = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < frames) list[atomicAdd(&starts[order_bucket(key[f], range)], 1u)] = (unsigned)f;
}

// the range starts as [UINT_MAX, 0]
__global__ void k_order_init(unsigned *counts, unsigned *range) {
    for (int i = threadIdx.x; i < ORDER_BUCKETS; i += blockDim.x) counts[i] = 0u;
    if (threadIdx.x == 0) {
        range[0] = 0xFFFFFFFFu;
        range[1] = 0u;
    }
}

size_t order_bytes(unsigned long long frames) { return ((size_t)frames * 2 + 2 * ORDER_BUCKETS + 2) * 4; }

hipError_t launch_order(const int8_t *in, int N, unsigned long long frames, void *scratch, hipStream_t s,
                        const unsigned **list_out) {
    if (frames == 0 || frames > 0xFFFFFFFFull) return hipErrorInvalidValue;
    unsigned *list = static_cast<unsigned *>(scratch);
    unsigned *key = list + frames;
    unsigned *counts = key + frames;
    unsigned *starts = counts + ORDER_BUCKETS;
    unsigned *range = starts + ORDER_BUCKETS;
    const dim3 g((unsigned)((frames + 255) / 256));
    hipLaunchKernelGGL(k_order_init, dim3(1), dim3(256), 0, s, counts, range);
    hipLaunchKernelGGL(k_order_keys, dim3((unsigned)((frames + 4 * ORDER_FPW - 1) / (4 * ORDER_FPW))), dim3(256), 0, s,
                       in, N, frames, key, range);
    hipLaunchKernelGGL(k_order_count, g, dim3(256), 0, s, key, frames, range, counts);
    hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(ORDER_BUCKETS), 0, s, counts, starts);
    hipLaunchKernelGGL(k_order_scatter, g, dim3(256), 0, s, key, frames, range, starts, list);
    *list_out = list;
    return hipGetLastError();
}

}  // namespace ldpc
