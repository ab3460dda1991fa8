// mixed_api.hip -- one decode step over a stream of frames from several codes.
//
// The caller hands over one byte buffer of frames of any of the codes, a code
// id and a byte offset per frame.  A bucket kernel splits the frame ids into
// one device queue per code (wave-aggregated atomics); each code's frame-queue
// kernel (its own autotuned words-per-workgroup, workgroup size and items in
// flight) then drains its queue on a stream of its own.  The kernels run side
// by side on the GPU: every one is launched with as many workgroups as stay
// resident, so workgroups of one code wait for CUs that another code's
// workgroups free when its queue runs dry.  The step is kernels and event
// edges only (queue counters are cleared by kernels), so it is captured once
// into a hipGraph and replayed while the buffers stay the same.
//
// ldpc_mixed_decode_noalloc runs the same step with its scratch (the per-code
// queue counters and frame lists) in `scratch`, a file-scope device array,
// instead of memory the library allocates, for callers that cannot allocate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ldpc_mi355x.h"
#include "code.h"
#include "kernels.h"
#include "pool.h"

namespace ldpc {
void set_last_error(const std::string &m);
}

namespace {

// frame f goes to queue cid[f]; a frame with an unknown code id is skipped
// and its iteration count set to 0
__global__ __launch_bounds__(256) void k_bucket(const uint8_t *__restrict__ cid, unsigned long long F, int n_codes,
                                                unsigned *counts, unsigned *lists, unsigned long long cap,
                                                int *iters) {
    const int lane = (int)__lane_id();
    for (unsigned long long f0 = (unsigned long long)blockIdx.x * blockDim.x; f0 < F;
         f0 += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long f = f0 + threadIdx.x;
        const int c = f < F ? (int)cid[f] : -1;
        for (int k = 0; k < n_codes; k++) {
            const unsigned long long m = __ballot(c == k);
            if (!m) continue;
            const int leader = __ffsll((long long)m) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&counts[k], (unsigned)__popcll(m));
            base = __shfl(base, leader);
            if (c == k) {
                const unsigned rank = (unsigned)__popcll(m & ((1ull << lane) - 1ull));
                lists[(unsigned long long)k * cap + base + rank] = (unsigned)f;
            }
        }
        if (f < F && c >= n_codes && iters) iters[f] = 0;
    }
}

// Scratch of ldpc_mixed_decode_noalloc: MIXED_MAX_CODES queue counters, then
// one frame list of `frames` entries per code.  Each device holds its own copy
// for the life of the process; host code uses the address that
// hipGetSymbolAddress returns for the current device, never the name.
constexpr size_t MIXED_MAX_CODES = 255;
constexpr size_t MIXED_COUNTS = 256;  // counters, padded to 1 KiB
static __device__ __align__(256) unsigned scratch[MIXED_COUNTS + LDPC_MIXED_NOALLOC_ENTRIES];
static_assert(MIXED_MAX_CODES <= MIXED_COUNTS, "one counter per code");

// the device address of `scratch` on the current device, looked up once per device
constexpr int MAX_DEVICES = 64;
unsigned *scratch_ptr[MAX_DEVICES] = {};
std::mutex scratch_mu;

int resolve_scratch(unsigned **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        ldpc::set_last_error(std::string("mixed scratch: no usable current device: ") + hipGetErrorString(e));
        return LDPC_ERR_HIP;
    }
    std::lock_guard<std::mutex> g(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            ldpc::set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return LDPC_ERR_HIP;
        }
        scratch_ptr[dev] = static_cast<unsigned *>(p);
    }
    *out = scratch_ptr[dev];
    return LDPC_OK;
}

struct GraphKey {
    const void *in = nullptr, *off = nullptr, *cid = nullptr, *out = nullptr, *iters = nullptr;
    uint64_t frames = 0;
    bool operator==(const GraphKey &o) const {
        return in == o.in && off == o.off && cid == o.cid && out == o.out && iters == o.iters && frames == o.frames;
    }
};

}  // namespace

struct ldpc_mixed {
    std::vector<ldpc_pool *> pools;
    std::vector<int> order;  // launch order: largest LDS block per workgroup first
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> done;
    hipEvent_t fork = nullptr;
    hipStream_t stream = nullptr;  // capture stream
    unsigned *d_counts = nullptr;
    unsigned *d_lists = nullptr;
    unsigned long long cap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    GraphKey key;
    int device = 0;
    unsigned *scratch = nullptr;  // device address of `scratch` on `device`
};

#define MHIP(expr)                                                                          \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            ldpc::set_last_error(std::string(#expr) + ": " + hipGetErrorString(_e));        \
            return LDPC_ERR_HIP;                                                            \
        }                                                                                   \
    } while (0)

namespace {

void drop_graph(ldpc_mixed *m) {
    if (m->exec) (void)hipGraphExecDestroy(m->exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    m->exec = nullptr;
    m->graph = nullptr;
    m->key = GraphKey();
}

void mixed_free(ldpc_mixed *m) {
    if (!m) return;
    drop_graph(m);
    for (ldpc_pool *p : m->pools) ldpc_pool_destroy(p);
    for (hipStream_t s : m->streams) (void)hipStreamDestroy(s);
    for (hipEvent_t e : m->done) (void)hipEventDestroy(e);
    if (m->fork) (void)hipEventDestroy(m->fork);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    (void)hipFree(m->d_counts);
    (void)hipFree(m->d_lists);
    delete m;
}

int ensure_capacity(ldpc_mixed *m, uint64_t frames) {
    if (frames <= m->cap) return LDPC_OK;
    MHIP(hipDeviceSynchronize());
    drop_graph(m);
    (void)hipFree(m->d_lists);
    m->d_lists = nullptr;
    m->cap = 0;
    MHIP(hipMalloc((void **)&m->d_lists, (size_t)frames * m->pools.size() * sizeof(unsigned)));
    m->cap = frames;
    return LDPC_OK;
}

// the step: clear counters, bucket, fork one frame-queue kernel per code, join.
// counts: one queue counter per code; lists: per code `cap` frame ids.
int enqueue_step(ldpc_mixed *m, unsigned *counts, unsigned *lists, unsigned long long cap, const int8_t *d_in,
                 const uint64_t *d_off, const uint8_t *d_cid, uint64_t frames, uint8_t *d_out, int *d_iters,
                 hipStream_t s) {
    const int nc = (int)m->pools.size();
    MHIP(ldpc::launch_zero_u32(counts, nc, s));
    // Largest LDS footprint per workgroup first: those workgroups need the
    // most contiguous room on a CU, and a kernel launched first fills the CUs
    // before the next queue starts dispatching (seen on graph replays); the
    // small-footprint kernels then fill the gaps.
    const unsigned grid = (unsigned)std::min<uint64_t>((frames + 255) / 256, 4096);
    hipLaunchKernelGGL(k_bucket, dim3(grid), dim3(256), 0, s, d_cid, (unsigned long long)frames, nc, counts, lists,
                       cap, d_iters);
    MHIP(hipGetLastError());
    MHIP(hipEventRecord(m->fork, s));
    for (int c : m->order) {
        MHIP(hipStreamWaitEvent(m->streams[c], m->fork, 0));
        ldpc::PoolIO io;
        io.in = d_in;
        io.out = d_out;
        io.iters = d_iters;
        io.total_ptr = counts + c;
        io.frame_list = lists + (size_t)c * cap;
        io.frame_off = reinterpret_cast<const unsigned long long *>(d_off);
        MHIP(ldpc::pool_enqueue(m->pools[c], io, m->pools[c]->max_grid, m->streams[c]));
        MHIP(hipEventRecord(m->done[c], m->streams[c]));
        MHIP(hipStreamWaitEvent(s, m->done[c], 0));
    }
    return LDPC_OK;
}

}  // namespace

extern "C" {

ldpc_mixed *ldpc_mixed_create(const ldpc_code *const *codes, int n_codes, int variant, int max_iterations,
                              int stop_rule, const ldpc_tuning *tunings, uint64_t max_frames) {
    if (!codes || n_codes < 1 || n_codes > 255 || max_iterations < 1) {
        ldpc::set_last_error("ldpc_mixed_create: bad arguments (1..255 codes)");
        return nullptr;
    }
    ldpc_mixed *m = new ldpc_mixed();
    for (int c = 0; c < n_codes; c++) {
        ldpc_pool *p = codes[c] ? ldpc_pool_create(codes[c], variant, max_iterations, stop_rule,
                                                   tunings ? &tunings[c] : nullptr)
                                : nullptr;
        if (!p) {
            const std::string why = codes[c] ? ldpc_last_error() : "null code";
            mixed_free(m);
            ldpc::set_last_error("ldpc_mixed_create: code " + std::to_string(c) + ": " + why);
            return nullptr;
        }
        m->pools.push_back(p);
    }
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    for (int c = 0; e == hipSuccess && c < n_codes; c++) {
        hipStream_t s = nullptr;
        hipEvent_t ev = nullptr;
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e == hipSuccess) m->streams.push_back(s);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) m->done.push_back(ev);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->fork, hipEventDisableTiming);
    for (int c = 0; c < n_codes; c++) m->order.push_back(c);
    auto lds = [m](int c) { return (size_t)m->pools[c]->code.N * m->pools[c]->P * (m->pools[c]->pairs ? 2 : 4); };
    std::stable_sort(m->order.begin(), m->order.end(), [&](int x, int y) { return lds(x) > lds(y); });
    if (e == hipSuccess) e = hipMalloc((void **)&m->d_counts, (size_t)n_codes * sizeof(unsigned));
    if (e != hipSuccess) {
        ldpc::set_last_error(std::string("ldpc_mixed_create: ") + hipGetErrorString(e));
        mixed_free(m);
        return nullptr;
    }
    if (max_frames && ensure_capacity(m, max_frames) != LDPC_OK) {
        mixed_free(m);
        return nullptr;
    }
    // set-up of the no-allocation step: this device's copy of `scratch`
    if (hipGetDevice(&m->device) != hipSuccess || resolve_scratch(&m->scratch) != LDPC_OK) {
        const std::string why = ldpc_last_error();
        mixed_free(m);
        ldpc::set_last_error("ldpc_mixed_create: " + why);
        return nullptr;
    }
    return m;
}

int ldpc_mixed_set_table(ldpc_mixed *m, int code_index, const int8_t *table, int n_iters, int n_degs) {
    if (!m || code_index < 0 || code_index >= (int)m->pools.size()) {
        ldpc::set_last_error("ldpc_mixed_set_table: bad arguments");
        return LDPC_ERR_ARG;
    }
    MHIP(hipDeviceSynchronize());
    drop_graph(m);  // a captured step holds the old table's address
    return ldpc_pool_set_table(m->pools[code_index], table, n_iters, n_degs);
}

int ldpc_mixed_decode(ldpc_mixed *m, const int8_t *d_in, const uint64_t *d_offsets, const uint8_t *d_code_ids,
                      uint64_t frames, uint8_t *d_out, int *d_iters, int use_graph, void *stream) {
    if (!m || !d_in || !d_offsets || !d_code_ids || !d_out || frames == 0 || frames > 0xFFFFFFFFull) {
        ldpc::set_last_error("ldpc_mixed_decode: bad arguments (1 <= frames < 2^32)");
        return LDPC_ERR_ARG;
    }
    for (ldpc_pool *p : m->pools)
        if ((p->variant == LDPC_LOMS || p->variant == LDPC_LNMS) && !p->d_table) {
            ldpc::set_last_error("learned variant needs a table for every code (ldpc_mixed_set_table)");
            return LDPC_ERR_STATE;
        }
    int rc = ensure_capacity(m, frames);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph)
        return enqueue_step(m, m->d_counts, m->d_lists, m->cap, d_in, d_offsets, d_code_ids, frames, d_out, d_iters, s);

    GraphKey k;
    k.in = d_in;
    k.off = d_offsets;
    k.cid = d_code_ids;
    k.out = d_out;
    k.iters = d_iters;
    k.frames = frames;
    if (!m->exec || !(k == m->key)) {
        drop_graph(m);
        MHIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_step(m, m->d_counts, m->d_lists, m->cap, d_in, d_offsets, d_code_ids, frames, d_out, d_iters,
                          m->stream);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(m->stream, &g);
        if (rc != LDPC_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) {
            ldpc::set_last_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
            return LDPC_ERR_HIP;
        }
        m->graph = g;
        const hipError_t ei = hipGraphInstantiate(&m->exec, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            drop_graph(m);
            ldpc::set_last_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
            return LDPC_ERR_HIP;
        }
        m->key = k;
    }
    MHIP(hipGraphLaunch(m->exec, s));
    return LDPC_OK;
}

int ldpc_mixed_decode_noalloc(ldpc_mixed *m, const int8_t *d_in, const uint64_t *d_offsets,
                              const uint8_t *d_code_ids, uint64_t frames, uint8_t *d_out, int *d_iters, void *stream) {
    if (!m || !d_in || !d_offsets || !d_code_ids || !d_out || frames == 0) {
        ldpc::set_last_error("ldpc_mixed_decode_noalloc: bad arguments");
        return LDPC_ERR_ARG;
    }
    const uint64_t nc = m->pools.size();
    if (frames > LDPC_MIXED_NOALLOC_ENTRIES / nc) {
        ldpc::set_last_error("ldpc_mixed_decode_noalloc: frames x codes = " + std::to_string(frames * nc) +
                             " exceeds the static scratch of " + std::to_string(LDPC_MIXED_NOALLOC_ENTRIES) +
                             " queue entries (LDPC_MIXED_NOALLOC_ENTRIES)");
        return LDPC_ERR_ARG;
    }
    for (ldpc_pool *p : m->pools)
        if ((p->variant == LDPC_LOMS || p->variant == LDPC_LNMS) && !p->d_table) {
            ldpc::set_last_error("learned variant needs a table for every code (ldpc_mixed_set_table)");
            return LDPC_ERR_STATE;
        }
    int dev = -1;
    MHIP(hipGetDevice(&dev));
    if (dev != m->device) {
        ldpc::set_last_error("ldpc_mixed_decode_noalloc: the current device is not the one the object was "
                             "created on");
        return LDPC_ERR_STATE;
    }
    return enqueue_step(m, m->scratch, m->scratch + MIXED_COUNTS, frames, d_in, d_offsets, d_code_ids, frames, d_out,
                        d_iters, (hipStream_t)stream);
}

int ldpc_mixed_config(const ldpc_mixed *m, int code_index, int *words_per_block, int *threads, int *max_workgroups,
                      int *uniform_deg, int *batch, int *pairs) {
    if (!m || code_index < 0 || code_index >= (int)m->pools.size()) {
        ldpc::set_last_error("ldpc_mixed_config: bad arguments");
        return LDPC_ERR_ARG;
    }
    return ldpc_pool_config(m->pools[code_index], words_per_block, threads, max_workgroups, uniform_deg, batch,
                            pairs);
}

void ldpc_mixed_destroy(ldpc_mixed *m) {
    if (!m) return;
    (void)hipDeviceSynchronize();
    mixed_free(m);
}

}  // extern "C"
