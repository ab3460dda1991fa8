// sim_kernels.hip -- device AWGN channel and device error / syndrome counting.
//
// Channel (replaces src/awgn_channel/CChanel_AWGN_SIMD.cu:7-112, which draws
// cuRAND XORWOW uniforms, applies Box-Muller and copies every chunk to the
// host): each sample s = frame * N + var is a pure function of (seed, snr
// index, s) -- one Philox4x32-10 block (Salmon et al., SC'11) gives two
// 53-bit uniforms, Box-Muller's cosine branch gives z, y = -1 + sigma * z
// (all-zero codeword, bit 0 -> -1 as in CChanel_AWGN_SIMD.cu:21-22), and the
// LLR is quantised on the device straight into the decoder's input layout.
//
// Error counting (replaces the host scan of src/ber_analyzer/CErrorAnalyzer.cpp:141-157):
// bit errors over the first n_info positions of each frame, frame errors, and a
// per-frame syndrome check against the decoder's parity-check table.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"

namespace ldpc {

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ int8_t channel_llr(uint64_t seed, uint32_t snr_idx, uint64_t s, double sigma,
                                              double llr_scale, int quant) {
    uint32_t c[4] = {(uint32_t)s, (uint32_t)(s >> 32), snr_idx, 0x4C445043u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t a = (((uint64_t)c[0] << 32) | c[1]) >> 11;
    const uint64_t b = (((uint64_t)c[2] << 32) | c[3]) >> 11;
    const double u1 = ((double)a + 1.0) * 0x1.0p-53;
    const double u2 = (double)b * 0x1.0p-53;
    const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
    const double y = -1.0 + sigma * z;
    if (quant == 1) {  // (char)fminf(fmaxf(8*v, -31), 31), CChanel_AWGN_SIMD.cu:23-24
        const float v = (float)y;
        return (int8_t)fminf(fmaxf(8.0f * v, -31.0f), 31.0f);
    }
    double q = round(llr_scale * y);  // 8 * 2y / sigma^2
    q = fmin(fmax(q, -127.0), 127.0);
    return (int8_t)q;
}

// frame-major: thread -> 4 consecutive variables of one frame (one u32 store)
__global__ __launch_bounds__(256) void k_channel_frames(int8_t *__restrict__ out, int N, size_t frames, uint64_t f0,
                                                        uint64_t seed, uint32_t snr_idx, double sigma,
                                                        double scale, int quant) {
    const size_t n4 = (size_t)(N + 3) / 4;
    const size_t total = frames * n4;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t f = i / n4;
        const int v0 = (int)(i % n4) * 4;
        const uint64_t s0 = (f0 + f) * (uint64_t)N;
        if ((N & 3) == 0) {
            uint32_t w = 0;
#pragma unroll
            for (int r = 0; r < 4; r++)
                w |= (uint32_t)(uint8_t)channel_llr(seed, snr_idx, s0 + v0 + r, sigma, scale, quant) << (8 * r);
            reinterpret_cast<uint32_t *>(out + f * (size_t)N)[v0 >> 2] = w;
        } else {
            for (int r = 0; r < 4 && v0 + r < N; r++)
                out[f * (size_t)N + v0 + r] = channel_llr(seed, snr_idx, s0 + v0 + r, sigma, scale, quant);
        }
    }
}

// Eb/N0 of frame f for the range channel: lo + (hi - lo) * u, u from its own
// Philox block (counter (f_lo, f_hi, snr_idx, 0x534E5252)).
__device__ __forceinline__ double frame_sigma(uint64_t seed, uint32_t snr_idx, uint64_t f, double lo, double hi,
                                              double rate) {
    uint32_t c[4] = {(uint32_t)f, (uint32_t)(f >> 32), snr_idx, 0x534E5252u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double ebn0 = lo + (hi - lo) * ((double)c[0] * 0x1.0p-32);
    return sqrt(0.5 / (rate * pow(10.0, ebn0 / 10.0)));
}

__global__ __launch_bounds__(256) void k_channel_range(int8_t *__restrict__ out, int N, size_t frames, uint64_t f0,
                                                       uint64_t seed, uint32_t snr_idx, double lo, double hi,
                                                       double rate, int quant) {
    const size_t n4 = (size_t)(N + 3) / 4;
    const size_t total = frames * n4;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t f = i / n4;
        const int v0 = (int)(i % n4) * 4;
        const double sigma = frame_sigma(seed, snr_idx, f0 + f, lo, hi, rate);
        const double scale = 16.0 / (sigma * sigma);
        const uint64_t s0 = (f0 + f) * (uint64_t)N;
        for (int r = 0; r < 4 && v0 + r < N; r++)
            out[f * (size_t)N + v0 + r] = channel_llr(seed, snr_idx, s0 + v0 + r, sigma, scale, quant);
    }
}

// device layout: word [v][w] holds frames 4w..4w+3 of variable v
__global__ __launch_bounds__(256) void k_channel_words(uint32_t *__restrict__ out, int N, size_t W, uint64_t f0,
                                                       uint64_t seed, uint32_t snr_idx, double sigma, double scale,
                                                       int quant) {
    const size_t total = (size_t)N * W;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t v = i / W, w = i % W;
        uint32_t x = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t s = (f0 + 4 * w + k) * (uint64_t)N + v;
            x |= (uint32_t)(uint8_t)channel_llr(seed, snr_idx, s, sigma, scale, quant) << (8 * k);
        }
        out[i] = x;
    }
}

// One workgroup per frame: bit errors over [0, n_info), frame error, and the
// syndrome over the table (checks in any order; a parity is order-free).
__global__ __launch_bounds__(256) void k_count(const uint8_t *__restrict__ hard, int N, size_t frames, int n_info,
                                               const uint2 *__restrict__ checks, const uint32_t *__restrict__ edge_var,
                                               int C, unsigned long long *__restrict__ counters,
                                               uint32_t *__restrict__ per_frame_bits, uint8_t *__restrict__ per_frame_ok) {
    extern __shared__ uint8_t s_hard[];
    __shared__ uint32_t s_err, s_bad;
    const int tid = threadIdx.x, T = blockDim.x;
    for (size_t f = blockIdx.x; f < frames; f += gridDim.x) {
        const uint8_t *h = hard + f * (size_t)N;
        if (tid == 0) {
            s_err = 0;
            s_bad = 0;
        }
        uint32_t e = 0;
        for (int v = tid; v < N; v += T) {
            const uint8_t x = h[v];
            s_hard[v] = x;
            e += (v < n_info) && (x != 0);
        }
        __syncthreads();
        if (e) atomicAdd(&s_err, e);
        uint32_t bad = 0;
        if (checks) {
            for (int c = tid; c < C; c += T) {
                const uint2 ci = checks[c];
                const int d = (int)(ci.x & 0xFFu);
                uint32_t par = 0;
                for (int j = 0; j < d; j++) par ^= s_hard[edge_var[ci.y + j]];
                bad |= par & 1u;
            }
            if (bad) atomicOr(&s_bad, 1u);
        }
        __syncthreads();
        if (tid == 0) {
            if (s_err) {
                atomicAdd(&counters[0], (unsigned long long)s_err);
                atomicAdd(&counters[1], 1ull);
            }
            if (checks && s_bad) atomicAdd(&counters[2], 1ull);
            if (per_frame_bits) per_frame_bits[f] = s_err;
            if (per_frame_ok) per_frame_ok[f] = checks ? (uint8_t)(s_bad == 0) : 1;
        }
        __syncthreads();
    }
}

hipError_t launch_channel_range(int8_t *out, int N, size_t frames, uint64_t f0, uint64_t seed, uint32_t snr_idx,
                                double lo_db, double hi_db, double rate, int quant, hipStream_t s) {
    const size_t items = frames * (size_t)((N + 3) / 4);
    const unsigned grid = (unsigned)std::min<size_t>((items + 255) / 256, 65536);
    hipLaunchKernelGGL(k_channel_range, dim3(grid), dim3(256), 0, s, out, N, frames, f0, seed, snr_idx, lo_db, hi_db,
                       rate, quant);
    return hipGetLastError();
}

hipError_t set_count_lds_limit() {
    static std::once_flag flag;
    static hipError_t result = hipSuccess;
    return once_per_process(flag, result, [] {
        return hipFuncSetAttribute((const void *)k_count, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(MAX_LDS_BYTES - 1024));
    });
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

hipError_t launch_channel(void *out, int layout, int N, size_t frames, uint64_t f0, uint64_t seed, uint32_t snr_idx,
                          double sigma, int quant, hipStream_t s) {
    const double scale = 16.0 / (sigma * sigma);
    if (layout == 0) {
        const size_t items = frames * (size_t)((N + 3) / 4);
        const unsigned grid = (unsigned)std::min<size_t>((items + 255) / 256, 65536);
        hipLaunchKernelGGL(k_channel_frames, dim3(grid), dim3(256), 0, s, (int8_t *)out, N, frames, f0, seed, snr_idx,
                           sigma, scale, quant);
    } else {
        const size_t W = frames / 4;
        const size_t items = (size_t)N * W;
        const unsigned grid = (unsigned)std::min<size_t>((items + 255) / 256, 65536);
        hipLaunchKernelGGL(k_channel_words, dim3(grid), dim3(256), 0, s, (uint32_t *)out, N, W, f0, seed, snr_idx,
                           sigma, scale, quant);
    }
    return hipGetLastError();
}

hipError_t launch_count(const uint8_t *hard, int N, size_t frames, int n_info, const uint2 *checks,
                        const uint32_t *edge_var, int C, unsigned long long *counters, uint32_t *per_frame_bits,
                        uint8_t *per_frame_ok, bool zero_first, hipStream_t s) {
    if (zero_first) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, counters, 3);
    const unsigned grid = (unsigned)std::min<size_t>(frames, 8192);
    hipLaunchKernelGGL(k_count, dim3(grid), dim3(256), (size_t)N, s, hard, N, frames, n_info, checks, edge_var, C,
                       counters, per_frame_bits, per_frame_ok);
    return hipGetLastError();
}

}  // namespace ldpc
