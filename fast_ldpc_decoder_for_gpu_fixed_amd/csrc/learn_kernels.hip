// learn_kernels.hip -- float shadow of the layered decoder, forward and
// reverse mode, for learning per-iteration / per-degree offset (or scale)
// tables that the integer decoder then uses as int8 steps.
//
// Forward (per frame, table order = dependency levels, in LSB units):
//     r = p - m (first iteration: p);  t = clamp(r, -128, 127)
//     mo_j = min of the other |t| (127 cap);  sign by the others' parity
//     mag = clamp(mo - beta[it][d], 0, 31)   or   clamp(alpha[it][d] * mo, 0, 31)
//     m_j = +-mag;  p = clamp(t + m_j, -128, 127)
//     loss += softplus(p / 8) / (T * N)  after every iteration
// Reverse: the gradient of a message reaches |t| of the edge holding min1
// (min2 for that edge itself), t through the clamps, and the table entry
// through the offset/scale; r values are stored by the forward pass and
// every other quantity is recomputed from them.
//
// Layout mirrors the integer kernel: a workgroup owns FB frames, their float
// posterior (or its gradient) lives in LDS as [var][FB]; per-edge floats
// (messages / message gradients, stored r) are [block][level-order check]
// [frame][j] so a lane's D values are contiguous.
#include <hip/hip_runtime.h>

#include "kernels.h"

#pragma clang fp contract(off)

namespace ldpc {

__device__ __forceinline__ float clampf_(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

template <int D>
struct CheckState {
    float t[D];
    float mn[D];
    float min1, min2;
    int i1, i2, par;
};

template <int D>
__device__ __forceinline__ void shadow_eval(const float (&r)[D], float pv, int mode, CheckState<D> &s) {
    s.min1 = 127.0f;
    s.min2 = 127.0f;
    s.i1 = -1;
    s.i2 = -1;
    s.par = 0;
#pragma unroll
    for (int j = 0; j < D; j++) {
        s.t[j] = clampf_(r[j], -128.0f, 127.0f);
        const float a = fabsf(s.t[j]);
        if (a < s.min1) {
            s.min2 = s.min1;
            s.i2 = s.i1;
            s.min1 = a;
            s.i1 = j;
        } else if (a < s.min2) {
            s.min2 = a;
            s.i2 = j;
        }
        s.par ^= s.t[j] > 0.0f;
    }
#pragma unroll
    for (int j = 0; j < D; j++) {
        const float mo = (j == s.i1) ? s.min2 : s.min1;
        const float mag = mode == 0 ? clampf_(mo - pv, 0.0f, 31.0f) : clampf_(pv * mo, 0.0f, 31.0f);
        s.mn[j] = (s.par ^ (s.t[j] > 0.0f)) ? mag : -mag;
    }
}

template <int D>
__device__ __forceinline__ void fwd_check(const ShadowArgs &a, float *post, float *mbase, float *rbase, int c, int f,
                                          bool first, float pv) {
    const uint2 ci = a.checks[c];
    const uint32_t *ev = a.edge_var + ci.y;
    float *mp = mbase + (size_t)ci.y * a.FB + (size_t)f * D;
    float *rp = rbase + (size_t)ci.y * a.FB + (size_t)f * D;
    uint32_t vi[D];
    float r[D];
#pragma unroll
    for (int j = 0; j < D; j++) vi[j] = (ev[j] << a.lfb) + (uint32_t)f;
#pragma unroll
    for (int j = 0; j < D; j++) r[j] = first ? post[vi[j]] : post[vi[j]] - mp[j];
#pragma unroll
    for (int j = 0; j < D; j++) rp[j] = r[j];
    CheckState<D> s;
    shadow_eval<D>(r, pv, a.mode, s);
#pragma unroll
    for (int j = 0; j < D; j++) {
        mp[j] = s.mn[j];
        post[vi[j]] = clampf_(s.t[j] + s.mn[j], -128.0f, 127.0f);
    }
}

template <int D>
__device__ __forceinline__ float bwd_check(const ShadowArgs &a, float *gpost, float *gmbase, const float *rbase,
                                           int c, int f, bool first, float pv) {
    const uint2 ci = a.checks[c];
    const uint32_t *ev = a.edge_var + ci.y;
    float *gmp = gmbase + (size_t)ci.y * a.FB + (size_t)f * D;
    const float *rp = rbase + (size_t)ci.y * a.FB + (size_t)f * D;
    uint32_t vi[D];
    float r[D], gt[D], ga[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
        vi[j] = (ev[j] << a.lfb) + (uint32_t)f;
        r[j] = rp[j];
        ga[j] = 0.0f;
    }
    CheckState<D> s;
    shadow_eval<D>(r, pv, a.mode, s);
    float gparam = 0.0f;
#pragma unroll
    for (int j = 0; j < D; j++) {
        const float u = s.t[j] + s.mn[j];
        const float gu = (u > -128.0f && u < 127.0f) ? gpost[vi[j]] : 0.0f;
        gt[j] = gu;
        const float gmn = gu + gmp[j];
        const float sg = (s.par ^ (s.t[j] > 0.0f)) ? 1.0f : -1.0f;
        const float gmag = sg * gmn;
        const float mo = (j == s.i1) ? s.min2 : s.min1;
        float gmo;
        if (a.mode == 0) {
            const float x = mo - pv;
            const bool inside = x > 0.0f && x < 31.0f;
            gmo = inside ? gmag : 0.0f;
            gparam -= gmo;
        } else {
            const float x = pv * mo;
            const bool inside = x > 0.0f && x < 31.0f;
            gmo = inside ? gmag * pv : 0.0f;
            gparam += inside ? gmag * mo : 0.0f;
        }
        const int dst = (j == s.i1) ? s.i2 : s.i1;
#pragma unroll
        for (int k = 0; k < D; k++) ga[k] += (k == dst) ? gmo : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < D; k++) {
        const float sgn = s.t[k] > 0.0f ? 1.0f : (s.t[k] < 0.0f ? -1.0f : 0.0f);
        gt[k] += ga[k] * sgn;
    }
#pragma unroll
    for (int j = 0; j < D; j++) {
        const bool inside = r[j] > -128.0f && r[j] < 127.0f;
        gpost[vi[j]] = inside ? gt[j] : 0.0f;
        gmp[j] = (!first && inside) ? -gt[j] : 0.0f;
    }
    return gparam;
}

#define LDPC_SHADOW_SWITCH(CALL)                                                            \
    switch (D) {                                                                            \
    case 1: CALL(1); break; case 2: CALL(2); break; case 3: CALL(3); break;                \
    case 4: CALL(4); break; case 5: CALL(5); break; case 6: CALL(6); break;                \
    case 7: CALL(7); break; case 8: CALL(8); break; case 9: CALL(9); break;                \
    case 10: CALL(10); break; case 11: CALL(11); break; case 12: CALL(12); break;          \
    case 13: CALL(13); break; case 14: CALL(14); break; case 15: CALL(15); break;          \
    case 16: CALL(16); break; default: break;                                               \
    }

__device__ __forceinline__ float softplus_(float x) { return fmaxf(x, 0.0f) + log1pf(expf(-fabsf(x))); }

__global__ __launch_bounds__(256) void k_shadow_forward(const ShadowArgs a) {
    extern __shared__ float s_post[];
    __shared__ float s_red[256];
    const int tid = threadIdx.x, T = blockDim.x, FB = a.FB, N = a.N;
    const size_t b = blockIdx.x;
    const int f = tid & (FB - 1);
    const int cstride = T >> a.lfb;
    const size_t nb = a.frames / FB;
    for (int i = tid; i < N * FB; i += T) {
        const int v = i >> a.lfb, fl = i & (FB - 1);
        s_post[i] = (float)a.llr[(b * FB + fl) * (size_t)N + v];
    }
    __syncthreads();
    float *mbase = a.msg + b * (size_t)a.E * FB;
    float lsum = 0.0f;
    const float wl = 1.0f / ((float)a.T * (float)N);
    for (int it = 0; it < a.T; it++) {
        const bool first = it == 0;
        float *rbase = a.rstore + ((size_t)it * nb + b) * (size_t)a.E * FB;
        for (int L = 0; L < a.n_levels; L++) {
            const int lb = a.level_beg[L], le = a.level_beg[L + 1];
            for (int c = lb + (tid >> a.lfb); c < le; c += cstride) {
                const int D = (int)(a.checks[c].x & 0xFFu);
                const float pv = a.params[(size_t)it * a.dcols + min(D, a.dcols - 1)];
#define LDPC_FWD(d) fwd_check<d>(a, s_post, mbase, rbase, c, f, first, pv)
                LDPC_SHADOW_SWITCH(LDPC_FWD)
#undef LDPC_FWD
            }
            __syncthreads();
        }
        float *pe = a.pend + ((size_t)it * nb + b) * (size_t)N * FB;
        for (int i = tid; i < N * FB; i += T) {
            const float p = s_post[i];
            pe[i] = p;
            lsum += softplus_(p / 8.0f) * wl;
        }
        __syncthreads();  // the dump reads LDS that the next iteration rewrites
    }
    if (a.post_out) {
        for (int i = tid; i < N * FB; i += T) {
            const int v = i >> a.lfb, fl = i & (FB - 1);
            a.post_out[(b * FB + fl) * (size_t)N + v] = s_post[i];
        }
    }
    s_red[tid] = lsum;
    __syncthreads();
    for (int s = T / 2; s > 0; s >>= 1) {
        if (tid < s) s_red[tid] += s_red[tid + s];
        __syncthreads();
    }
    if (tid == 0) atomicAdd(a.loss, s_red[0]);
}

__global__ __launch_bounds__(256) void k_shadow_backward(const ShadowArgs a) {
    extern __shared__ float s_gp[];
    __shared__ float s_grad[MAX_DEG + 1];
    const int tid = threadIdx.x, T = blockDim.x, FB = a.FB, N = a.N;
    const size_t b = blockIdx.x;
    const int f = tid & (FB - 1);
    const int cstride = T >> a.lfb;
    const size_t nb = a.frames / FB;
    for (int i = tid; i < N * FB; i += T) s_gp[i] = 0.0f;
    float *gmbase = a.msg + b * (size_t)a.E * FB;  // message buffer reused for message gradients
    for (size_t i = tid; i < (size_t)a.E * FB; i += T) gmbase[i] = 0.0f;
    const float wl = 1.0f / ((float)a.T * (float)N);
    for (int it = a.T - 1; it >= 0; it--) {
        const bool first = it == 0;
        if (tid <= MAX_DEG) s_grad[tid] = 0.0f;
        __syncthreads();
        const float *pe = a.pend + ((size_t)it * nb + b) * (size_t)N * FB;
        for (int i = tid; i < N * FB; i += T) {
            const float x = pe[i] / 8.0f;
            s_gp[i] += wl * (1.0f / (1.0f + expf(-x))) / 8.0f;
        }
        __syncthreads();
        const float *rbase = a.rstore + ((size_t)it * nb + b) * (size_t)a.E * FB;
        for (int L = a.n_levels - 1; L >= 0; L--) {
            const int lb = a.level_beg[L], le = a.level_beg[L + 1];
            for (int c = lb + (tid >> a.lfb); c < le; c += cstride) {
                const int D = (int)(a.checks[c].x & 0xFFu);
                const int col = min(D, a.dcols - 1);
                const float pv = a.params[(size_t)it * a.dcols + col];
                float g = 0.0f;
#define LDPC_BWD(d) g = bwd_check<d>(a, s_gp, gmbase, rbase, c, f, first, pv)
                LDPC_SHADOW_SWITCH(LDPC_BWD)
#undef LDPC_BWD
                atomicAdd(&s_grad[col], g);
            }
            __syncthreads();
        }
        if (tid < a.dcols) atomicAdd(&a.grad[(size_t)it * a.dcols + tid], s_grad[tid]);
        __syncthreads();
    }
}

hipError_t launch_shadow(const ShadowArgs &a, bool backward, hipStream_t s) {
    const size_t lds = (size_t)a.N * a.FB * sizeof(float);
    static std::once_flag flag;
    static hipError_t attr = hipSuccess;
    const hipError_t e = once_per_process(flag, attr, [] {
        hipError_t r = hipFuncSetAttribute((const void *)k_shadow_forward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(MAX_LDS_BYTES - 2048));
        if (r == hipSuccess)
            r = hipFuncSetAttribute((const void *)k_shadow_backward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(MAX_LDS_BYTES - 2048));
        return r;
    });
    if (e != hipSuccess) return e;
    const unsigned nb = (unsigned)(a.frames / a.FB);
    if (backward)
        hipLaunchKernelGGL(k_shadow_backward, dim3(nb), dim3(256), lds, s, a);
    else
        hipLaunchKernelGGL(k_shadow_forward, dim3(nb), dim3(256), lds, s, a);
    return hipGetLastError();
}

}  // namespace ldpc
