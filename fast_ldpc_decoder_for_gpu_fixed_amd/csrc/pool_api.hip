// pool_api.hip -- C ABI of the persistent frame-queue decoder.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ldpc_mi355x.h"
#include "code.h"
#include "kernels.h"

namespace ldpc {
void set_last_error(const std::string &m);
}

struct ldpc_pool {
    ldpc_code code;
    ldpc::Plan plan;
    int variant = 0, max_iters = 0, P = 1, log2P = 0, threads = 256, max_grid = 1;
    hipStream_t stream = nullptr;
    uint2 *d_checks = nullptr;
    uint32_t *d_edge_var = nullptr;
    int *d_level_beg = nullptr;
    uint32_t *d_msgs = nullptr;
    unsigned long long *d_next = nullptr;
    int8_t *d_table = nullptr;
    int tits = 0, tdeg = 0;
};

#define PHIP(expr)                                                                          \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            ldpc::set_last_error(std::string(#expr) + ": " + hipGetErrorString(_e));        \
            return LDPC_ERR_HIP;                                                            \
        }                                                                                   \
    } while (0)

namespace {
void pool_free(ldpc_pool *p) {
    if (!p) return;
    hipFree(p->d_checks);
    hipFree(p->d_edge_var);
    hipFree(p->d_level_beg);
    hipFree(p->d_msgs);
    hipFree(p->d_next);
    hipFree(p->d_table);
    if (p->stream) hipStreamDestroy(p->stream);
    delete p;
}

__global__ void k_zero_counter(unsigned long long *p) { *p = 0ull; }

ldpc::PoolArgs pool_args(const ldpc_pool *p) {
    ldpc::PoolArgs a;
    std::memset(&a, 0, sizeof a);
    a.checks = p->d_checks;
    a.edge_var = p->d_edge_var;
    a.level_beg = p->d_level_beg;
    a.n_levels = p->plan.n_levels;
    a.N = p->code.N;
    a.C = p->code.C;
    a.E = p->code.E;
    a.P = p->P;
    a.log2P = p->log2P;
    a.next = p->d_next;
    a.msgs = p->d_msgs;
    a.max_iters = p->max_iters;
    a.variant = p->variant;
    a.ltable = p->d_table;
    a.tits = p->tits;
    a.tdeg = p->tdeg;
    return a;
}
}  // namespace

extern "C" {

ldpc_pool *ldpc_pool_create(const ldpc_code *code, int variant, int max_iterations, const ldpc_tuning *tuning) {
    if (!code || max_iterations < 1 || variant < LDPC_MS || variant > LDPC_LNMS) {
        ldpc::set_last_error("ldpc_pool_create: bad arguments");
        return nullptr;
    }
    if (variant == LDPC_OMS && code->q1_split < code->C) {
        ldpc::set_last_error("OMS with two degree groups has a first-iteration rule that is not per frame; "
                             "use LDPC_OMS_V2 (clamp everywhere)");
        return nullptr;
    }
    if (code->max_deg > ldpc::MAX_DEG) {
        ldpc::set_last_error("check degree above 32");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        ldpc::set_last_error("no HIP device visible");
        return nullptr;
    }
    ldpc_tuning tu = {0, 0, -1, -1, -1, 0};
    if (tuning) tu = *tuning;
    const size_t budget = ldpc::MAX_LDS_BYTES - 4096;
    const size_t wb = (size_t)code->N * 4;
    int P = tu.words_per_block;
    if (P <= 0) {
        P = 1;
        while (P * 2 <= ldpc::MAX_P && wb * (size_t)(P * 2) <= 40 * 1024) P *= 2;
    }
    if (P < 1 || P > ldpc::MAX_P || (P & (P - 1)) || wb * (size_t)P > budget) {
        ldpc::set_last_error("words_per_block must be a power of two <= 16 whose posterior block fits LDS");
        return nullptr;
    }
    int T = tu.threads > 0 ? tu.threads : 256;
    if (T % 64 || T % P || T > ldpc::max_threads_for_degree(code->max_deg, false) || T < 4 * P) {
        ldpc::set_last_error("threads must be a multiple of 64 and of words_per_block, >= 4*words_per_block, "
                             "within the kernel limit");
        return nullptr;
    }
    ldpc_pool *p = new ldpc_pool();
    p->code = *code;
    p->plan = ldpc::make_plan(p->code);
    p->variant = variant;
    p->max_iters = max_iterations;
    p->P = P;
    while ((1 << p->log2P) < P) p->log2P++;
    p->threads = T;
    hipDeviceProp_t prop;
    int dev = 0;
    hipGetDevice(&dev);
    hipGetDeviceProperties(&prop, dev);
    ldpc::PoolArgs a = pool_args(p);
    p->max_grid = prop.multiProcessorCount * ldpc::pool_blocks_per_cu(a, code->max_deg, T);
    std::vector<uint2> ck((size_t)code->C);
    for (int i = 0; i < code->C; i++) ck[i] = make_uint2(p->plan.check_info[i], p->plan.check_e0[i]);
    hipError_t e = hipSuccess;
    auto al = [&](auto **q, size_t n) {
        if (e == hipSuccess && n) e = hipMalloc((void **)q, n * sizeof(**q));
    };
    al(&p->d_checks, ck.size());
    al(&p->d_edge_var, (size_t)code->E);
    al(&p->d_level_beg, p->plan.level_beg.size());
    al(&p->d_msgs, (size_t)p->max_grid * code->E * P);
    al(&p->d_next, 1);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->stream, hipStreamDefault);
    if (e == hipSuccess) e = hipMemcpy(p->d_checks, ck.data(), ck.size() * sizeof(uint2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_edge_var, p->plan.edge_var.data(), (size_t)code->E * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_level_beg, p->plan.level_beg.data(), p->plan.level_beg.size() * sizeof(int),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ldpc::set_last_error(std::string("ldpc_pool_create: ") + hipGetErrorString(e));
        pool_free(p);
        return nullptr;
    }
    return p;
}

int ldpc_pool_set_table(ldpc_pool *p, const int8_t *table, int n_iters, int n_degs) {
    if (!p || !table || n_iters < 1 || n_degs < 1 || n_iters > 64 || n_degs > ldpc::MAX_DEG + 1) {
        ldpc::set_last_error("ldpc_pool_set_table: bad table (at most 64 rows, 33 columns)");
        return LDPC_ERR_ARG;
    }
    const int hi = p->variant == LDPC_LNMS ? 8 : 127;
    for (int i = 0; i < n_iters * n_degs; i++)
        if (table[i] < 0 || table[i] > hi) {
            ldpc::set_last_error("table value out of range");
            return LDPC_ERR_ARG;
        }
    PHIP(hipStreamSynchronize(p->stream));
    hipFree(p->d_table);
    p->d_table = nullptr;
    PHIP(hipMalloc((void **)&p->d_table, (size_t)n_iters * n_degs));
    PHIP(hipMemcpy(p->d_table, table, (size_t)n_iters * n_degs, hipMemcpyHostToDevice));
    p->tits = n_iters;
    p->tdeg = n_degs;
    return LDPC_OK;
}

int ldpc_pool_decode(ldpc_pool *p, const int8_t *d_in, uint64_t frames, uint8_t *d_out, int *d_iters, int workgroups,
                     void *stream) {
    if (!p || !d_in || !d_out || frames == 0) {
        ldpc::set_last_error("ldpc_pool_decode: bad arguments");
        return LDPC_ERR_ARG;
    }
    if ((p->variant == LDPC_LOMS || p->variant == LDPC_LNMS) && !p->d_table) {
        ldpc::set_last_error("learned variant needs ldpc_pool_set_table first");
        return LDPC_ERR_STATE;
    }
    hipStream_t s = stream ? (hipStream_t)stream : p->stream;
    const uint64_t slots = 4ull * (uint64_t)p->P;
    int grid = workgroups > 0 ? std::min(workgroups, p->max_grid) : p->max_grid;
    grid = (int)std::min<uint64_t>((uint64_t)grid, (frames + slots - 1) / slots);
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, s, p->d_next);
    PHIP(hipGetLastError());
    ldpc::PoolArgs a = pool_args(p);
    a.in_frames = d_in;
    a.out_hard = d_out;
    a.out_iters = d_iters;
    a.total = frames;
    PHIP(ldpc::launch_pool(a, p->code.max_deg, grid, p->threads, s));
    return LDPC_OK;
}

int ldpc_pool_config(const ldpc_pool *p, int *words_per_block, int *threads, int *max_workgroups) {
    if (!p) {
        ldpc::set_last_error("null pool");
        return LDPC_ERR_ARG;
    }
    if (words_per_block) *words_per_block = p->P;
    if (threads) *threads = p->threads;
    if (max_workgroups) *max_workgroups = p->max_grid;
    return LDPC_OK;
}

void *ldpc_pool_stream(ldpc_pool *p) { return p ? (void *)p->stream : nullptr; }

void ldpc_pool_destroy(ldpc_pool *p) {
    if (p) hipStreamSynchronize(p->stream);
    pool_free(p);
}

}  // extern "C"
