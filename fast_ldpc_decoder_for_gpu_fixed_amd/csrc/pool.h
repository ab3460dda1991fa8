// pool.h -- the frame-queue decoder object, shared by its C ABI and the
// mixed-code decoder built from one such object per code.
#pragma once
#include <hip/hip_runtime.h>

#include "code.h"
#include "kernels.h"

struct ldpc_pool {
    ldpc_code code;
    ldpc::Plan plan;
    int variant = 0, max_iters = 0, P = 1, log2P = 0, threads = 256, max_grid = 1;
    hipStream_t stream = nullptr;
    uint2 *d_checks = nullptr;
    uint32_t *d_edge_var = nullptr;
    uint32_t *d_edge_lds = nullptr;  // LDS byte address of each edge's variable (the kernels' gathers)
    int *d_level_beg = nullptr;
    uint32_t *d_msgs = nullptr;
    unsigned long long *d_next = nullptr;
    int8_t *d_table = nullptr;
    int tits = 0, tdeg = 0;
    uint32_t *d_edge_off16 = nullptr;  // single-degree path
    int uniform_deg = 0, batch = 1;
    int stop_rule = 0;
    int pairs = 0;  // posterior and messages as 2-frame pairs (P counts pairs)
    uint32_t *d_rows = nullptr;  // pair mode with degrees <= 8: LDS addresses as rows of 8 per check
    int *d_level_deg = nullptr;  // degree shared by every check of a level, or 0
    int n_cu = 1;   // compute units of the device
    int order = 0;  // LDPC_ORDER_*: queue order of the frames of a launch
    void *d_order = nullptr;  // its scratch (ldpc::order_bytes of the largest batch so far)
    size_t order_cap = 0;
    // captured decode step (ldpc_pool_decode_graph) and what it was captured with
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    struct {
        const void *in, *out, *iters, *table, *order;
        unsigned long long frames;
        int grid, max_iters;
    } key = {};
};


namespace ldpc {
// where one launch takes its frames from and writes them to
struct PoolIO {
    const int8_t *in = nullptr;
    uint8_t *out = nullptr;
    int *iters = nullptr;
    unsigned long long total = 0;
    const unsigned *total_ptr = nullptr;
    const unsigned *frame_list = nullptr;
    const unsigned long long *frame_off = nullptr;
};
PoolArgs pool_args(const ldpc_pool *p);
// zero the pool's queue head and launch its kernel on s (capturable: kernels
// only).  With the pool's order set and no frame list in io, the order
// kernels run first and the launch takes its frames through their list: the
// order scratch must already hold order_bytes(io.total) (pool_reserve_order).
hipError_t pool_enqueue(ldpc_pool *p, const PoolIO &io, int grid, hipStream_t s);
// grow the pool's order scratch to `frames` frames (allocates: not under capture)
hipError_t pool_reserve_order(ldpc_pool *p, unsigned long long frames);
}  // namespace ldpc
