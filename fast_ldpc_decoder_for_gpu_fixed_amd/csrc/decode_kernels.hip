// decode_kernels.hip -- layered 8-bit Min-Sum decoding on gfx950.
//
// What is computed (per 32-bit word = 4 frames, per check, in the serial
// table order of the reference; src/decoder_oms/cuda/CUDA_OMS_SIMD.cu:148-191):
//     t_j  = sat8(post[v_j] - msg_j)            (first iteration: post[v_j])
//     min1, min2 = two smallest |t_j| (unsigned, |-128| = 128, both start at 127)
//     S    = parity of {t_j > 0}
//     msg_j = +-(|t_j| == min1 ? f(min2) : g(min1)),  sign = parity of the others
//     post[v_j] = sat8(t_j + msg_j)
//
// How it is laid out on the chip:
//   * one workgroup owns a block of P words (4P frames) for the whole decode;
//     its posterior lives in LDS as [var][P] words;
//   * the table is cut into levels (no two checks of a level share a
//     variable, and every check comes after all earlier-table checks it
//     shares a variable with), so a level's checks run in parallel and the
//     result is bit-identical to the serial order; a workgroup barrier
//     separates levels;
//   * lane (check, word) reads its D messages as one contiguous run
//     (messages are stored [block][level-order edge][word][j]) and gathers /
//     scatters its D posterior words in LDS;
//   * int8 lanes are computed two frames per packed-16 op (byte_lanes.h).
#include "check_core.h"

namespace ldpc {

template <int D, bool EQ, typename PostPtr>
__device__ __forceinline__ uint32_t check_d(PostPtr post, const uint32_t *__restrict__ ev,
                                            uint32_t *__restrict__ mp, int lp, int w, bool first,
                                            int variant, bool nocap, int lparam) {
    uint32_t vi[D], pw[D], mw[D], pn[D], mn[D];
#pragma unroll
    for (int j = 0; j < D; j++) vi[j] = (ev[j] << lp) + (uint32_t)w;  // P is a power of two
#pragma unroll
    for (int j = 0; j < D; j++) pw[j] = post[vi[j]];
    if (!first) {
#pragma unroll
        for (int j = 0; j < D; j++) mw[j] = mp[j];
    }
    const u16x2 lpv = splat(learned_half(variant, lparam));
    const uint2 pr = check_core<D, EQ>(pw, mw, first, variant, nocap, lpv, lpv, pn, mn);
    const uint32_t parity = pr.x | pr.y;
#pragma unroll
    for (int j = 0; j < D; j++) {
        mp[j] = mn[j];
        post[vi[j]] = pn[j];
    }
    return parity;
}

// One level's items of one thread for a code whose checks all have degree D
// (the shared body is uniform_items in check_core.h).
template <int D, int G, bool EQ, bool FIRST, int VK>
__device__ __forceinline__ uint32_t level_uniform(const DecodeArgs &a, uint32_t *post, uint32_t *mbase,
                                                  int c_first, int le, int cstride, int w, int lparam) {
    const u16x2 lpv = splat(learned_half(a.variant, lparam));
    const uint2 pr = uniform_items<D, G, EQ, FIRST, VK>(a.edge_off16, a.P, a.variant, post, mbase, c_first, le, cstride,
                                                    w, 0xFFFFFFFFu, lpv, lpv);
    return pr.x | pr.y;
}

template <bool EQ, int DMAXC, typename PostPtr>
__device__ __forceinline__ uint32_t check_any(int D, PostPtr post, const uint32_t *__restrict__ ev,
                                              uint32_t *__restrict__ mp, int lp, int w, bool first,
                                              int variant, bool nocap, int lparam) {
    switch (D) {
#define LDPC_CASE(d)                                                                        \
    case d:                                                                                 \
        if constexpr (d <= DMAXC)                                                           \
            return check_d<d, EQ>(post, ev, mp, lp, w, first, variant, nocap, lparam);      \
        else                                                                                \
            return 0;
        LDPC_CASE(1) LDPC_CASE(2) LDPC_CASE(3) LDPC_CASE(4) LDPC_CASE(5) LDPC_CASE(6)
        LDPC_CASE(7) LDPC_CASE(8) LDPC_CASE(9) LDPC_CASE(10) LDPC_CASE(11) LDPC_CASE(12)
        LDPC_CASE(13) LDPC_CASE(14) LDPC_CASE(15) LDPC_CASE(16) LDPC_CASE(17) LDPC_CASE(18)
        LDPC_CASE(19) LDPC_CASE(20) LDPC_CASE(21) LDPC_CASE(22) LDPC_CASE(23) LDPC_CASE(24)
        LDPC_CASE(25) LDPC_CASE(26) LDPC_CASE(27) LDPC_CASE(28) LDPC_CASE(29) LDPC_CASE(30)
        LDPC_CASE(31) LDPC_CASE(32)
#undef LDPC_CASE
    default: return 0;
    }
}

// ---------------------------------------------------------------------------
// posterior block load / store (LDS [var][P] <-> the three global layouts)
// ---------------------------------------------------------------------------
__device__ void load_block(const DecodeArgs &a, uint32_t *post, size_t b) {
    const int T = blockDim.x, tid = threadIdx.x, P = a.P, N = a.N;
    if (a.load_mode == LOAD_FRAMES) {
        // frames [4bP, 4bP+4P) are contiguous: 4P*N bytes
        const int8_t *src = a.in_frames + b * (size_t)(4 * P) * (size_t)N;
        uint8_t *lb = reinterpret_cast<uint8_t *>(post);
        if ((N & 3) == 0) {
            const int n4 = N >> 2;
            for (int fl = 0; fl < 4 * P; fl++) {
                const uint32_t *s32 = reinterpret_cast<const uint32_t *>(src + (size_t)fl * N);
                const int wl = fl >> 2, k = fl & 3;
                for (int q = tid; q < n4; q += T) {
                    const uint32_t x = s32[q];
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        lb[((size_t)(4 * q + r) * P + wl) * 4 + k] = (uint8_t)(x >> (8 * r));
                }
            }
        } else {
            for (int fl = 0; fl < 4 * P; fl++) {
                const int wl = fl >> 2, k = fl & 3;
                for (int v = tid; v < N; v += T)
                    lb[((size_t)v * P + wl) * 4 + k] = (uint8_t)src[(size_t)fl * N + v];
            }
        }
    } else if (a.load_mode == LOAD_WORDS) {
        const int lp = a.log2P;
        for (int i = tid; i < N * P; i += T) {
            const int v = i >> lp, wl = i & (P - 1);
            post[i] = a.in_words[(size_t)v * a.W + b * P + wl];
        }
    } else {
        const uint32_t *src = a.post_state + b * (size_t)N * P;
        for (int i = tid; i < N * P; i += T) post[i] = src[i];
    }
}

__device__ void store_block(const DecodeArgs &a, const uint32_t *post, size_t b) {
    const int T = blockDim.x, tid = threadIdx.x, P = a.P, N = a.N;
    if (a.store_state) {
        uint32_t *dst = a.post_state + b * (size_t)N * P;
        for (int i = tid; i < N * P; i += T) dst[i] = post[i];
    }
    if (a.out_hard) {
        uint8_t *dst = a.out_hard + b * (size_t)(4 * P) * (size_t)N;
        const uint8_t *lb = reinterpret_cast<const uint8_t *>(post);
        if ((N & 3) == 0) {
            const int n4 = N >> 2;
            for (int fl = 0; fl < 4 * P; fl++) {
                uint32_t *d32 = reinterpret_cast<uint32_t *>(dst + (size_t)fl * N);
                const int wl = fl >> 2, k = fl & 3;
                for (int q = tid; q < n4; q += T) {
                    uint32_t x = 0;
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int8_t s = (int8_t)lb[((size_t)(4 * q + r) * P + wl) * 4 + k];
                        x |= (uint32_t)(s > 0) << (8 * r);  // GPU_Transpose_uint8.cu:29
                    }
                    d32[q] = x;
                }
            }
        } else {
            for (int fl = 0; fl < 4 * P; fl++) {
                const int wl = fl >> 2, k = fl & 3;
                for (int v = tid; v < N; v += T) {
                    const int8_t s = (int8_t)lb[((size_t)v * P + wl) * 4 + k];
                    dst[(size_t)fl * N + v] = (uint8_t)(s > 0);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// the decode kernel: one workgroup per block of P words, all iterations
// ---------------------------------------------------------------------------
// The generic path for degrees <= 16 is held to 168 VGPRs (3 waves per SIMD;
// with the posterior in global memory it needs 169 unconstrained).
template <bool EQ, bool POST_IN_LDS, int DMAXC, int UD, int G, int VK>
__global__ __launch_bounds__(UD > 0 ? 512 : (DMAXC <= 8 ? 1024 : 256))
__attribute__((amdgpu_waves_per_eu(UD == 0 && DMAXC == 16 ? 3 : 1))) void k_decode(const DecodeArgs a) {
    extern __shared__ uint32_t lds_post[];
    __shared__ uint32_t s_ov[MAX_P];
    __shared__ int s_active[MAX_P];
    __shared__ int8_t s_lrow[MAX_DEG + 1];

    const int tid = threadIdx.x, T = blockDim.x, P = a.P, lp = a.log2P;
    const size_t b = a.block0 + blockIdx.x;
    const int w = tid & (P - 1);
    uint32_t *post = POST_IN_LDS ? lds_post : a.post_state + b * (size_t)a.N * P;

    if (POST_IN_LDS || a.load_mode != LOAD_STATE) {
        if (POST_IN_LDS) {
            load_block(a, post, b);
        } else {
            // global-posterior path: bring the input into the state block
            const int N = a.N;
            if (a.load_mode == LOAD_WORDS) {
                for (int i = tid; i < N * P; i += T) {
                    const int v = i >> lp, wl = i & (P - 1);
                    post[i] = a.in_words[(size_t)v * a.W + b * P + wl];
                }
            } else {
                const int8_t *src = a.in_frames + b * (size_t)(4 * P) * (size_t)N;
                uint8_t *pb = reinterpret_cast<uint8_t *>(post);
                for (int fl = 0; fl < 4 * P; fl++) {
                    const int wl = fl >> 2, k = fl & 3;
                    for (int v = tid; v < N; v += T)
                        pb[((size_t)v * P + wl) * 4 + k] = (uint8_t)src[(size_t)fl * N + v];
                }
            }
        }
    }
    if (tid < P) s_active[tid] = a.stopped ? (a.stopped[b * P + tid] == 0) : 1;
    __syncthreads();

    uint32_t *mbase = a.msgs + b * (size_t)a.E * P;
    const int cstride = T >> lp;
    int ran = 0;
    for (int it = 0; it < a.iters; it++) {
        const bool first = a.first && it == 0;
        if (a.ltable) {
            const int row = min(a.iter_base + it, a.tits - 1);
            for (int d = tid; d <= MAX_DEG; d += T)
                s_lrow[d] = a.ltable[(size_t)row * a.tdeg + min(d, a.tdeg - 1)];
        }
        if (tid < P) s_ov[tid] = 0;
        __syncthreads();
        const bool act = s_active[w] != 0;
        uint32_t ov = 0;
        int lb = a.level_beg[0], le = a.level_beg[1];
        for (int L = 0; L < a.n_levels; L++) {
            const int ln = a.level_beg[min(L + 2, a.n_levels)];  // the next level's end, loaded early
            if constexpr (UD > 0) {
                if (act) {
                    const int lpar = a.ltable ? (int)s_lrow[UD] : 0;
                    if (first)
                        ov |= level_uniform<UD, G, EQ, true, VK>(a, post, mbase, lb + (tid >> lp), le, cstride, w, lpar);
                    else
                        ov |= level_uniform<UD, G, EQ, false, VK>(a, post, mbase, lb + (tid >> lp), le, cstride, w, lpar);
                }
            } else if (act) {
                for (int c = lb + (tid >> lp); c < le; c += cstride) {
                    const uint2 ci = a.checks[c];  // x: degree | q1 << 8, y: first edge
                    const int D = (int)(ci.x & 0xFFu);
                    const bool nocap = first && (ci.x & 0x100u) && a.variant == VAR_OMS;
                    const int lparam = a.ltable ? (int)s_lrow[D] : 0;
                    ov |= check_any<EQ, DMAXC>(D, post, a.edge_var + ci.y,
                                        mbase + (size_t)ci.y * P + (size_t)w * D, lp, w, first,
                                        a.variant, nocap, lparam);
                }
            }
            __syncthreads();
            lb = le;
            le = ln;
        }
        if (act) ran++;
        if (a.early_term && !first) {
            if (act && ov) atomicOr(&s_ov[w], 1u);
            __syncthreads();
            if (tid < P && s_active[tid] && s_ov[tid] == 0) {
                s_active[tid] = 0;
                if (a.stopped) a.stopped[b * P + tid] = 1;
            }
            __syncthreads();
            int any = 0;
            for (int q = 0; q < P; q++) any |= s_active[q];
            if (!any) break;
        }
    }
    if (a.iters_run && tid < P) a.iters_run[b * P + tid] += ran;  // tid < P => w == tid

    if (POST_IN_LDS) {
        store_block(a, post, b);
    } else if (a.out_hard) {
        __syncthreads();
        DecodeArgs c = a;
        c.store_state = 0;
        store_block(c, post, b);
    }
}

// ---------------------------------------------------------------------------
// device interleave (src/transpose/GPU_Transpose_uint8.cu:80-130) and inverse
// with hard decision (:9-78), as LDS-tiled transposes: a workgroup moves a
// tile of TV variables x TW words; global reads and writes are contiguous.
// Word tiles run along grid x (up to 2^31 - 1 of them), variable tiles
// along grid y (N up to 65535 * TV).
// ---------------------------------------------------------------------------
constexpr int TV = 64, TW = 64;

__global__ __launch_bounds__(256) void k_interleave(const int8_t *__restrict__ in,
                                                    uint32_t *__restrict__ out, int N, int W) {
    __shared__ uint32_t tile[TV][TW + 1];
    const int v0 = blockIdx.y * TV, w0 = blockIdx.x * TW, tid = threadIdx.x;
    uint8_t *tb = reinterpret_cast<uint8_t *>(&tile[0][0]);
    const int rowb = (TW + 1) * 4;
    // read 4*TW frames x TV bytes
    for (int i = tid; i < 4 * TW * TV; i += 256) {
        const int fl = i / TV, vl = i % TV;
        const int v = v0 + vl, wl = fl >> 2;
        if (v < N && w0 + wl < W)
            tb[vl * rowb + fl] = (uint8_t)in[((size_t)w0 * 4 + fl) * N + v];
    }
    __syncthreads();
    for (int i = tid; i < TV * TW; i += 256) {
        const int vl = i / TW, wl = i % TW;
        if (v0 + vl < N && w0 + wl < W) out[(size_t)(v0 + vl) * W + w0 + wl] = tile[vl][wl];
    }
}

// HD: the hard decision of each byte (the reference's inverse interleave);
// else the bytes themselves (device words back to frame-major LLRs)
template <bool HD>
__global__ __launch_bounds__(256) void k_deinterleave(const uint32_t *__restrict__ in,
                                                      uint8_t *__restrict__ out, int N, int W) {
    __shared__ uint32_t tile[TV][TW + 1];
    const int v0 = blockIdx.y * TV, w0 = blockIdx.x * TW, tid = threadIdx.x;
    for (int i = tid; i < TV * TW; i += 256) {
        const int vl = i / TW, wl = i % TW;
        uint32_t x = 0;
        if (v0 + vl < N && w0 + wl < W) x = in[(size_t)(v0 + vl) * W + w0 + wl];
        if constexpr (HD) {
            // vsetgts4(x, 0): byte -> 1 if > 0 (GPU_Transpose_uint8.cu:29)
            uint32_t hd = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) hd |= (uint32_t)((int8_t)(x >> (8 * k)) > 0) << (8 * k);
            x = hd;
        }
        tile[vl][wl] = x;
    }
    __syncthreads();
    const uint8_t *tb = reinterpret_cast<const uint8_t *>(&tile[0][0]);
    const int rowb = (TW + 1) * 4;
    for (int i = tid; i < 4 * TW * TV; i += 256) {
        const int fl = i / TV, vl = i % TV;
        const int v = v0 + vl, wl = fl >> 2;
        if (v < N && w0 + wl < W) out[((size_t)w0 * 4 + fl) * N + v] = tb[vl * rowb + fl];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <bool EQ, bool LDS, int DMAXC, int UD, int G, int VK = -1>
static hipError_t launch_decode_t(const DecodeArgs &a, int threads, hipStream_t s) {
    const size_t lds = LDS ? (size_t)a.N * a.P * 4 : 0;
    static std::once_flag flag;
    static hipError_t attr = hipSuccess;
    const hipError_t e = once_per_process(flag, attr, [] {
        return hipFuncSetAttribute((const void *)k_decode<EQ, LDS, DMAXC, UD, G, VK>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(MAX_LDS_BYTES - 1024));  // static flags live in the rest
    });
    if (e != hipSuccess) return e;
    if (a.block0 + a.nblocks > a.W / a.P) return hipErrorInvalidValue;
    const unsigned nblk = a.nblocks ? a.nblocks : (unsigned)(a.W / a.P - a.block0);
    hipLaunchKernelGGL((k_decode<EQ, LDS, DMAXC, UD, G, VK>), dim3(nblk), dim3(threads), lds, s, a);
    return hipGetLastError();
}

template <bool EQ, bool LDS>
static hipError_t launch_dclass(const DecodeArgs &a, int dmax, int threads, hipStream_t s) {
    if (dmax <= 8) return launch_decode_t<EQ, LDS, 8, 0, 1>(a, threads, s);
    if (dmax <= 16) return launch_decode_t<EQ, LDS, 16, 0, 1>(a, threads, s);
    return launch_decode_t<EQ, LDS, 32, 0, 1>(a, threads, s);
}

template <bool EQ, int UD>
static hipError_t launch_uniform_g(const DecodeArgs &a, int batch, int threads, hipStream_t s) {
    // the plain OMS family and learned offsets get their own straight-line
    // instantiations at every batch size; MS, NMS and 2NMS at the default
    // batch of 2 items; the rest read the variant at run time
    const int vk = (a.variant == VAR_OMS || a.variant == VAR_OMS_V2) ? VAR_OMS : a.variant;
#define LDPC_UG(VKC)                                                                              \
    if (vk == VKC) {                                                                              \
        if (batch >= 4) return launch_decode_t<EQ, true, 8, UD, 4, VKC>(a, threads, s);           \
        if (batch == 2) return launch_decode_t<EQ, true, 8, UD, 2, VKC>(a, threads, s);           \
        return launch_decode_t<EQ, true, 8, UD, 1, VKC>(a, threads, s);                           \
    }
    if constexpr (!EQ) {
        LDPC_UG(VAR_OMS)
        LDPC_UG(VAR_LOMS)
        if (batch == 2 && vk == VAR_MS) return launch_decode_t<EQ, true, 8, UD, 2, VAR_MS>(a, threads, s);
        if (batch == 2 && vk == VAR_NMS) return launch_decode_t<EQ, true, 8, UD, 2, VAR_NMS>(a, threads, s);
    } else {
        if (batch == 2 && vk == VAR_2NMS) return launch_decode_t<EQ, true, 8, UD, 2, VAR_2NMS>(a, threads, s);
    }
#undef LDPC_UG
    if (batch >= 4) return launch_decode_t<EQ, true, 8, UD, 4, -1>(a, threads, s);
    if (batch == 2) return launch_decode_t<EQ, true, 8, UD, 2, -1>(a, threads, s);
    return launch_decode_t<EQ, true, 8, UD, 1, -1>(a, threads, s);
}

template <bool EQ>
static hipError_t launch_uniform(const DecodeArgs &a, int ud, int batch, int threads, hipStream_t s) {
    switch (ud) {
    case 6: return launch_uniform_g<EQ, 6>(a, batch, threads, s);
    case 7: return launch_uniform_g<EQ, 7>(a, batch, threads, s);
    case 8: return launch_uniform_g<EQ, 8>(a, batch, threads, s);
    default: return hipErrorInvalidValue;
    }
}

bool uniform_degree_supported(int d) { return d >= 6 && d <= 8; }

int max_threads_for_degree(int dmax, bool uniform) { return uniform ? 512 : (dmax <= 8 ? 1024 : 256); }

hipError_t launch_decode(const DecodeArgs &a, int dmax, int threads, bool post_in_lds, hipStream_t s,
                         int uniform_deg, int batch) {
    if (dmax < 1 || dmax > MAX_DEG || threads > max_threads_for_degree(dmax, uniform_deg > 0)) return hipErrorInvalidValue;
    const bool eq = a.variant == VAR_2NMS;
    if (uniform_deg > 0) {
        if (!post_in_lds || !a.edge_off16 || uniform_deg != dmax) return hipErrorInvalidValue;
        return eq ? launch_uniform<true>(a, uniform_deg, batch, threads, s)
                  : launch_uniform<false>(a, uniform_deg, batch, threads, s);
    }
    if (post_in_lds) return eq ? launch_dclass<true, true>(a, dmax, threads, s) : launch_dclass<false, true>(a, dmax, threads, s);
    return eq ? launch_dclass<true, false>(a, dmax, threads, s) : launch_dclass<false, false>(a, dmax, threads, s);
}

hipError_t launch_interleave(const int8_t *in, uint32_t *out, int N, int W, hipStream_t s) {
    dim3 g((unsigned)((W + TW - 1) / TW), (unsigned)((N + TV - 1) / TV));
    hipLaunchKernelGGL(k_interleave, g, dim3(256), 0, s, in, out, N, W);
    return hipGetLastError();
}

hipError_t launch_deinterleave_hd(const uint32_t *in, uint8_t *out, int N, int W, hipStream_t s) {
    dim3 g((unsigned)((W + TW - 1) / TW), (unsigned)((N + TV - 1) / TV));
    hipLaunchKernelGGL(k_deinterleave<true>, g, dim3(256), 0, s, in, out, N, W);
    return hipGetLastError();
}

hipError_t launch_deinterleave_raw(const uint32_t *in, int8_t *out, int N, int W, hipStream_t s) {
    dim3 g((unsigned)((W + TW - 1) / TW), (unsigned)((N + TV - 1) / TV));
    hipLaunchKernelGGL(k_deinterleave<false>, g, dim3(256), 0, s, in, reinterpret_cast<uint8_t *>(out), N, W);
    return hipGetLastError();
}

}  // namespace ldpc
