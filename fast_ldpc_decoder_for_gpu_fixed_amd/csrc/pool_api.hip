// pool_api.hip -- C ABI of the persistent frame-queue decoder.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ldpc_mi355x.h"
#include "code.h"
#include "kernels.h"
#include "pool.h"

namespace ldpc {
void set_last_error(const std::string &m);
}

#define PHIP(expr)                                                                          \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            ldpc::set_last_error(std::string(#expr) + ": " + hipGetErrorString(_e));        \
            return LDPC_ERR_HIP;                                                            \
        }                                                                                   \
    } while (0)

namespace {
void drop_graph(ldpc_pool *p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    p->key = {};
}

void pool_free(ldpc_pool *p) {
    if (!p) return;
    drop_graph(p);
    if (p->capture) (void)hipStreamDestroy(p->capture);
    (void)hipFree(p->d_checks);
    (void)hipFree(p->d_edge_var);
    (void)hipFree(p->d_edge_lds);
    (void)hipFree(p->d_level_beg);
    (void)hipFree(p->d_msgs);
    (void)hipFree(p->d_next);
    (void)hipFree(p->d_table);
    (void)hipFree(p->d_edge_off16);
    (void)hipFree(p->d_rows);
    (void)hipFree(p->d_level_deg);
    (void)hipFree(p->d_order);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

__global__ void k_zero_counter(unsigned long long *p) { *p = 0ull; }
}  // namespace

namespace ldpc {
PoolArgs pool_args(const ldpc_pool *p) {
    ldpc::PoolArgs a;
    std::memset(&a, 0, sizeof a);
    a.checks = p->d_checks;
    a.edge_var = p->d_edge_var;
    a.edge_lds = p->d_edge_lds;
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
    a.edge_off16 = p->d_edge_off16;
    a.uniform_deg = p->uniform_deg;
    a.batch = p->batch;
    a.stop_rule = p->stop_rule;
    a.pairs = p->pairs;
    a.rows = p->d_rows;
    a.level_deg = p->d_level_deg;
    return a;
}

hipError_t pool_enqueue(ldpc_pool *p, const PoolIO &io, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, s, p->d_next);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    PoolArgs a = pool_args(p);
    a.in_frames = io.in;
    a.out_hard = io.out;
    a.out_iters = io.iters;
    a.total = io.total;
    a.total_ptr = io.total_ptr;
    a.frame_list = io.frame_list;
    a.frame_off = io.frame_off;
    if (p->order == LDPC_ORDER_LEAST_RELIABLE_FIRST && !io.frame_list && !io.frame_off && !io.total_ptr) {
        if (p->order_cap < order_bytes(io.total)) return hipErrorInvalidValue;
        e = launch_order(io.in, p->code.N, io.total, p->d_order, s, &a.frame_list);
        if (e != hipSuccess) return e;
    }
    return launch_pool(a, p->code.max_deg, grid, p->threads, s);
}

hipError_t pool_reserve_order(ldpc_pool *p, unsigned long long frames) {
    if (frames > 0xFFFFFFFFull) return hipErrorInvalidValue;  // queue positions are 32-bit frame ids
    const size_t need = order_bytes(frames);
    if (p->order_cap >= need) return hipSuccess;
    drop_graph(p);  // a captured step holds the old scratch's address
    (void)hipFree(p->d_order);
    p->d_order = nullptr;
    p->order_cap = 0;
    const hipError_t e = hipMalloc(&p->d_order, need);
    if (e == hipSuccess) p->order_cap = need;
    return e;
}
}  // namespace ldpc

extern "C" {

ldpc_pool *ldpc_pool_create(const ldpc_code *code, int variant, int max_iterations, int stop_rule,
                            const ldpc_tuning *tuning) {
    if (!code || max_iterations < 1 || variant < LDPC_MS || variant > LDPC_LNMS ||
        (stop_rule != LDPC_STOP_SYNDROME && stop_rule != LDPC_STOP_V2C_PARITY && stop_rule != LDPC_STOP_NONE)) {
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
    // posterior format: 4-frame words, or 2-frame pairs when a word block
    // does not fit LDS (tu.post_in_lds: -1 auto, 1 words, 2 pairs)
    if (tu.post_in_lds != -1 && tu.post_in_lds != 1 && tu.post_in_lds != 2) {
        ldpc::set_last_error("post_in_lds for the frame-queue decoder: -1 auto, 1 words of 4 frames, 2 pairs");
        return nullptr;
    }
    const bool words_fit = (size_t)code->N * 4 <= ldpc::POOL_LDS_DYNAMIC;
    const bool pairs = tu.post_in_lds == 2 || (tu.post_in_lds == -1 && !words_fit);
    const size_t wb = (size_t)code->N * (pairs ? 2 : 4);  // LDS bytes per word / pair
    const int fpw = pairs ? 2 : 4;                         // frames per word / pair
    int P = tu.words_per_block;
    if (P <= 0) {
        P = 1;
        while (P * 2 <= ldpc::MAX_P && wb * (size_t)(P * 2) <= 40 * 1024) P *= 2;
    }
    if (P < 1 || P > ldpc::MAX_P || (P & (P - 1)) || wb * (size_t)P > ldpc::POOL_LDS_DYNAMIC) {
        ldpc::set_last_error("words_per_block must be a power of two <= 16 whose posterior block fits LDS "
                             "(N * 4 * P bytes, or N * 2 * P in pair mode)");
        return nullptr;
    }
    // single-degree path: every check of degree 6..8 and LDS offsets within 16 bits
    int dmin = ldpc::MAX_DEG + 1;
    for (int x : code->deg) dmin = std::min(dmin, x);
    const bool uni_ok = !pairs && dmin == code->max_deg && ldpc::uniform_degree_supported(dmin) &&
                        wb * (size_t)P <= 65535;
    if (tu.uniform == 1 && !uni_ok) {
        ldpc::set_last_error("single-degree path requested but the code/blocking does not qualify");
        return nullptr;
    }
    const int ud = (tu.uniform != 0 && uni_ok) ? dmin : 0;
    const int batch = tu.batch > 0 ? tu.batch : 2;
    if (batch != 1 && batch != 2 && batch != 4) {
        ldpc::set_last_error("batch must be 1, 2 or 4");
        return nullptr;
    }
    const int tmax = ldpc::max_threads_for_degree(code->max_deg, ud > 0);
    int T = tu.threads > 0 ? tu.threads : std::min(pairs ? 1024 : 256, tmax);
    if (T % 64 || T % P || T > tmax || T < fpw * P) {
        ldpc::set_last_error("threads must be a multiple of 64 and of words_per_block, >= the frame slots, "
                             "at most " + std::to_string(tmax) + " for this code");
        return nullptr;
    }
    ldpc_pool *p = new ldpc_pool();
    p->code = *code;
    // word blocks: level sizes in multiples of 64 / P checks (full waves); pair mode keeps the
    // dependency levels, whose single-degree levels take the straight-line path
    p->plan = ldpc::make_plan(p->code, pairs ? 1 : 64 / P);
    p->variant = variant;
    p->max_iters = max_iterations;
    p->stop_rule = stop_rule;
    p->pairs = pairs ? 1 : 0;
    p->P = P;
    while ((1 << p->log2P) < P) p->log2P++;
    p->threads = T;
    p->uniform_deg = ud;
    p->batch = ud ? batch : 1;
    std::vector<uint32_t> off16;
    if (ud) {
        const int H = (ud + 1) / 2;
        off16.assign((size_t)code->C * H, 0);
        for (int cs = 0; cs < code->C; cs++)
            for (int j = 0; j < ud; j++) {
                const uint32_t o = (p->plan.edge_var[(size_t)cs * ud + j] << p->log2P) * 4;  // LDS byte offset
                off16[(size_t)cs * H + j / 2] |= (j & 1) ? (o << 16) : o;
            }
    }
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        ldpc::set_last_error("ldpc_pool_create: cannot query the current device");
        pool_free(p);
        return nullptr;
    }
    const ldpc::PoolArgs a = ldpc::pool_args(p);
    const int per_cu = ldpc::pool_prepare(a, code->max_deg, T);
    if (per_cu < 1) {
        ldpc::set_last_error(per_cu < 0 ? "ldpc_pool_create: the pair-mode kernel has static LDS in front of its "
                                          "posterior block (its rows address LDS from 0)"
                                        : "the frame-queue kernel cannot run with this words_per_block / threads");
        pool_free(p);
        return nullptr;
    }
    p->n_cu = prop.multiProcessorCount;
    p->max_grid = p->n_cu * per_cu;
    std::vector<uint2> ck((size_t)code->C);
    for (int i = 0; i < code->C; i++) ck[i] = make_uint2(p->plan.check_info[i], p->plan.check_e0[i]);
    hipError_t e = hipSuccess;
    auto al = [&](auto **q, size_t n) {
        if (e == hipSuccess && n) e = hipMalloc((void **)q, n * sizeof(**q));
    };
    // pair mode with degrees <= 8: variables as rows of 8 LDS byte offsets per
    // check (variable * 2P, the first pair of the variable's group), messages
    // as runs of 8 u16 per check and pair (16-byte aligned, addressed by the check)
    // LDS byte address of each edge's variable (its first word or pair)
    const uint32_t vstride = (p->pairs ? 2u : 4u) * (uint32_t)P;
    std::vector<uint32_t> edge_lds((size_t)code->E + ldpc::GATHER_PAD, 0u);
    for (size_t i = 0; i < (size_t)code->E; i++) edge_lds[i] = p->plan.edge_var[i] * vstride;
    std::vector<uint32_t> rows;
    if (p->pairs && code->max_deg <= 8) {
        rows.assign((size_t)code->C * 8, (uint32_t)code->N * vstride);  // unused entries: the padding variable
        for (int c = 0; c < code->C; c++) {
            const int D = (int)(p->plan.check_info[c] & 0xFFu);
            for (int j = 0; j < D; j++)
                rows[(size_t)c * 8 + j] = p->plan.edge_var[p->plan.check_e0[c] + j] * vstride;
        }
    }
    const size_t msg_u16 = rows.empty() ? (size_t)code->E : (size_t)code->C * 8;  // per frame group
    al(&p->d_checks, ck.size());
    al(&p->d_edge_var, (size_t)code->E + ldpc::GATHER_PAD);
    al(&p->d_edge_lds, edge_lds.size());
    al(&p->d_level_beg, p->plan.level_beg.size());
    al(&p->d_level_deg, p->plan.level_deg.size());
    al(&p->d_msgs, ((size_t)p->max_grid * msg_u16 * P * wb / code->N + 3) / 4 + ldpc::GATHER_PAD);  // u32 units
    al(&p->d_rows, rows.size());
    al(&p->d_next, 1);
    al(&p->d_edge_off16, off16.size());
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->stream, hipStreamDefault);
    if (e == hipSuccess) e = hipMemcpy(p->d_checks, ck.data(), ck.size() * sizeof(uint2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_edge_var, p->plan.edge_var.data(), (size_t)code->E * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(p->d_edge_var + code->E, 0, (size_t)ldpc::GATHER_PAD * 4);
    if (e == hipSuccess) e = hipMemcpy(p->d_edge_lds, edge_lds.data(), edge_lds.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_level_beg, p->plan.level_beg.data(), p->plan.level_beg.size() * sizeof(int),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_level_deg, p->plan.level_deg.data(), p->plan.level_deg.size() * sizeof(int),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess && !off16.empty())
        e = hipMemcpy(p->d_edge_off16, off16.data(), off16.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && !rows.empty()) e = hipMemcpy(p->d_rows, rows.data(), rows.size() * 4, hipMemcpyHostToDevice);
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
    drop_graph(p);  // a captured step holds the old table's address
    (void)hipFree(p->d_table);
    p->d_table = nullptr;
    PHIP(hipMalloc((void **)&p->d_table, (size_t)n_iters * n_degs));
    PHIP(hipMemcpy(p->d_table, table, (size_t)n_iters * n_degs, hipMemcpyHostToDevice));
    p->tits = n_iters;
    p->tdeg = n_degs;
    // the table takes LDS next to the posterior block: fewer workgroups may
    // stay resident (the message buffer was sized for the larger count)
    const int per_cu = ldpc::pool_prepare(ldpc::pool_args(p), p->code.max_deg, p->threads);
    if (per_cu < 1) {
        ldpc::set_last_error("the frame-queue kernel cannot run with this table");
        return LDPC_ERR_ARG;
    }
    p->max_grid = std::min(p->max_grid, p->n_cu * per_cu);
    return LDPC_OK;
}

int ldpc_pool_set_order(ldpc_pool *p, int order) {
    if (!p || (order != LDPC_ORDER_INPUT && order != LDPC_ORDER_LEAST_RELIABLE_FIRST)) {
        ldpc::set_last_error("ldpc_pool_set_order: bad arguments");
        return LDPC_ERR_ARG;
    }
    if (order != p->order) drop_graph(p);  // a captured step holds the old order's kernels
    p->order = order;
    return LDPC_OK;
}

static int pool_decode_args(ldpc_pool *p, const int8_t *d_in, uint64_t frames, uint8_t *d_out, int workgroups,
                            const char *who, int *grid) {
    if (!p || !d_in || !d_out || frames == 0) {
        ldpc::set_last_error(std::string(who) + ": bad arguments");
        return LDPC_ERR_ARG;
    }
    if ((p->variant == LDPC_LOMS || p->variant == LDPC_LNMS) && !p->d_table) {
        ldpc::set_last_error("learned variant needs ldpc_pool_set_table first");
        return LDPC_ERR_STATE;
    }
    const uint64_t slots = (p->pairs ? 2ull : 4ull) * (uint64_t)p->P;
    const int g = workgroups > 0 ? std::min(workgroups, p->max_grid) : p->max_grid;
    *grid = (int)std::min<uint64_t>((uint64_t)g, (frames + slots - 1) / slots);
    return LDPC_OK;
}

int ldpc_pool_decode(ldpc_pool *p, const int8_t *d_in, uint64_t frames, uint8_t *d_out, int *d_iters, int workgroups,
                     void *stream) {
    int grid = 0;
    const int rc = pool_decode_args(p, d_in, frames, d_out, workgroups, "ldpc_pool_decode", &grid);
    if (rc) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : p->stream;
    if (p->order) PHIP(ldpc::pool_reserve_order(p, frames));
    ldpc::PoolIO io;
    io.in = d_in;
    io.out = d_out;
    io.iters = d_iters;
    io.total = frames;
    PHIP(ldpc::pool_enqueue(p, io, grid, s));
    return LDPC_OK;
}

int ldpc_pool_decode_graph(ldpc_pool *p, const int8_t *d_in, uint64_t frames, uint8_t *d_out, int *d_iters,
                           int workgroups, void *stream) {
    int grid = 0;
    int rc = pool_decode_args(p, d_in, frames, d_out, workgroups, "ldpc_pool_decode_graph", &grid);
    if (rc) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : p->stream;
    if (p->order) PHIP(ldpc::pool_reserve_order(p, frames));
    const void *order = p->order ? p->d_order : nullptr;
    const bool same = p->exec && p->key.in == d_in && p->key.out == d_out && p->key.iters == d_iters &&
                      p->key.table == p->d_table && p->key.frames == frames && p->key.grid == grid &&
                      p->key.max_iters == p->max_iters && p->key.order == order;
    if (!same) {
        // capture once: the queue-head reset kernel and the decode kernel
        // (kernels only, no memset node), on the pool's capture stream
        drop_graph(p);
        if (!p->capture) PHIP(hipStreamCreateWithFlags(&p->capture, hipStreamNonBlocking));
        PHIP(hipStreamBeginCapture(p->capture, hipStreamCaptureModeThreadLocal));
        ldpc::PoolIO io;
        io.in = d_in;
        io.out = d_out;
        io.iters = d_iters;
        io.total = frames;
        const hipError_t el = ldpc::pool_enqueue(p, io, grid, p->capture);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(p->capture, &g);
        if (el != hipSuccess || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            ldpc::set_last_error(std::string("ldpc_pool_decode_graph: capture: ") +
                                 hipGetErrorString(el != hipSuccess ? el : ec));
            return LDPC_ERR_HIP;
        }
        p->graph = g;
        const hipError_t ei = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            drop_graph(p);
            ldpc::set_last_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
            return LDPC_ERR_HIP;
        }
        p->key.in = d_in;
        p->key.out = d_out;
        p->key.iters = d_iters;
        p->key.table = p->d_table;
        p->key.frames = frames;
        p->key.grid = grid;
        p->key.max_iters = p->max_iters;
        p->key.order = order;
    }
    PHIP(hipGraphLaunch(p->exec, s));
    return LDPC_OK;
}

int ldpc_pool_simulate(ldpc_pool *p, double ebn0_db, int quant, uint64_t seed, uint32_t snr_idx, uint64_t n_frames,
                       uint64_t batch_frames, ldpc_sim_stats *out) {
    if (!p || !out || n_frames == 0 || (quant != 0 && quant != 1)) {
        ldpc::set_last_error("ldpc_pool_simulate: bad arguments");
        return LDPC_ERR_ARG;
    }
    const ldpc_code &c = p->code;
    if ((size_t)c.N > ldpc::MAX_LDS_BYTES - 1024) {
        ldpc::set_last_error("frame longer than the counting kernel's LDS staging buffer");
        return LDPC_ERR_ARG;
    }
    const uint64_t batch = std::min<uint64_t>(n_frames, batch_frames ? batch_frames : (1ull << 18));
    const double rate = (double)(c.N - c.C) / (double)c.N;
    const double sigma = ldpc_channel_sigma(ebn0_db, rate, 0);
    PHIP(ldpc::set_count_lds_limit());  // once per process
    int8_t *d_llr = nullptr;
    uint8_t *d_hard = nullptr;
    int *d_it = nullptr;
    unsigned long long *d_cnt = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LDPC_OK;
    double ms = 0.0, it_sum = 0.0;
    std::vector<int> its((size_t)batch);
    hipStream_t s = p->stream;
    hipError_t e = hipMalloc((void **)&d_llr, (size_t)batch * c.N);
    if (e == hipSuccess) e = hipMalloc((void **)&d_hard, (size_t)batch * c.N);
    if (e == hipSuccess) e = hipMalloc((void **)&d_it, (size_t)batch * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&d_cnt, 3 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (uint64_t f0 = 0; e == hipSuccess && rc == LDPC_OK && f0 < n_frames; f0 += batch) {
        const uint64_t nb = std::min<uint64_t>(batch, n_frames - f0);
        e = ldpc::launch_channel(d_llr, 0, c.N, nb, f0, seed, snr_idx, sigma, quant, s);
        if (e != hipSuccess) break;
        e = hipEventRecord(e0, s);
        if (e != hipSuccess) break;
        rc = ldpc_pool_decode(p, d_llr, nb, d_hard, d_it, 0, s);
        if (rc) break;
        e = hipEventRecord(e1, s);
        if (e == hipSuccess)
            e = ldpc::launch_count(d_hard, c.N, nb, c.N - c.C, p->d_checks, p->d_edge_var, c.C, d_cnt, nullptr,
                                   nullptr, f0 == 0, s);
        if (e == hipSuccess) e = hipMemcpyAsync(its.data(), d_it, nb * sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) break;
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        ms += t;
        for (uint64_t i = 0; i < nb; i++) it_sum += its[i];
    }
    unsigned long long cnt[3] = {0, 0, 0};
    if (e == hipSuccess && rc == LDPC_OK) e = hipMemcpy(cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost);
    if (e != hipSuccess && rc == LDPC_OK) {
        ldpc::set_last_error(std::string("ldpc_pool_simulate: ") + hipGetErrorString(e));
        rc = LDPC_ERR_HIP;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_llr);
    (void)hipFree(d_hard);
    (void)hipFree(d_it);
    (void)hipFree(d_cnt);
    if (rc) return rc;
    out->frames = n_frames;
    out->bit_errors = cnt[0];
    out->frame_errors = cnt[1];
    out->syndrome_failures = cnt[2];
    out->avg_iterations = it_sum / (double)n_frames;
    out->decode_ms = ms;
    return LDPC_OK;
}

int ldpc_pool_autotune(const ldpc_code *code, int variant, int max_iterations, int stop_rule, double ebn0_db,
                       int quant, uint64_t sample_frames, ldpc_tuning *best, double *best_ms, int *n_tried) {
    if (!code || !best || max_iterations < 1 || sample_frames == 0 || (quant != 0 && quant != 1)) {
        ldpc::set_last_error("ldpc_pool_autotune: bad arguments");
        return LDPC_ERR_ARG;
    }
    const double rate = (double)(code->N - code->C) / (double)code->N;
    const double sigma = ldpc_channel_sigma(ebn0_db, rate, 0);
    int8_t *d_in = nullptr;
    uint8_t *d_out = nullptr;
    int *d_it = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void **)&d_in, (size_t)sample_frames * code->N);
    if (e == hipSuccess) e = hipMalloc((void **)&d_out, (size_t)sample_frames * code->N);
    if (e == hipSuccess) e = hipMalloc((void **)&d_it, (size_t)sample_frames * sizeof(int));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = ldpc::launch_channel(d_in, 0, code->N, sample_frames, 0, 0x7475E5EEDull, 0, sigma, quant, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    std::vector<int8_t> table((size_t)max_iterations * (ldpc::MAX_DEG + 1), variant == LDPC_LNMS ? 6 : 1);
    double best_t = 1e300;
    int tried = 0;
    std::string last = "no configuration could run";
    for (int fmt = 1; fmt <= 2; fmt++)
    for (int P = 1; e == hipSuccess && P <= ldpc::MAX_P; P *= 2) {
        if ((size_t)code->N * (fmt == 1 ? 4 : 2) * P > ldpc::POOL_LDS_DYNAMIC) break;
        // workgroup sizes in whole waves, not only powers of two: a level of
        // n checks fills n * P / T of a round's lanes (QC 12x24 Z=81: 81
        // checks per level, 162 items at P = 2: 63% of 256 lanes, 84% of 192)
        for (int T : {64, 128, 192, 256, 320, 384, 512, 768, 1024}) {
            if (e != hipSuccess) break;
            if (T % P || T < (fmt == 1 ? 4 : 2) * P) continue;
            for (int uni = fmt == 1 ? 1 : 0; uni >= 0; uni--) {
                for (int batch : {1, 2, 4}) {
                    if (!uni && batch != 1) continue;
                    ldpc_tuning tu = {P, T, -1, fmt, uni, batch};
                    ldpc_pool *pool = ldpc_pool_create(code, variant, max_iterations, stop_rule, &tu);
                    if (!pool) continue;  // this combination does not qualify
                    if (uni && !pool->uniform_deg) {
                        ldpc_pool_destroy(pool);
                        continue;
                    }
                    int rc = LDPC_OK;
                    if (variant == LDPC_LOMS || variant == LDPC_LNMS)
                        rc = ldpc_pool_set_table(pool, table.data(), max_iterations, ldpc::MAX_DEG + 1);
                    float t_min = 1e30f;
                    for (int rep = 0; rc == LDPC_OK && rep < 3; rep++) {
                        (void)hipEventRecord(e0, pool->stream);
                        rc = ldpc_pool_decode(pool, d_in, sample_frames, d_out, d_it, 0, nullptr);
                        (void)hipEventRecord(e1, pool->stream);
                        if (rc == LDPC_OK && hipEventSynchronize(e1) == hipSuccess) {
                            float t = 0.f;
                            (void)hipEventElapsedTime(&t, e0, e1);
                            if (rep > 0) t_min = std::min(t_min, t);  // the first run warms up
                        }
                    }
                    ldpc_pool_destroy(pool);
                    if (rc != LDPC_OK) {
                        last = ldpc_last_error();
                        continue;
                    }
                    tried++;
                    if (t_min < best_t) {
                        best_t = t_min;
                        *best = tu;
                        best->batch = uni ? batch : 0;
                    }
                }
                if (!uni) break;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    (void)hipFree(d_it);
    if (e != hipSuccess) {
        ldpc::set_last_error(std::string("ldpc_pool_autotune: ") + hipGetErrorString(e));
        return LDPC_ERR_HIP;
    }
    if (!tried) {
        ldpc::set_last_error("ldpc_pool_autotune: " + last);
        return LDPC_ERR_ARG;
    }
    if (best_ms) *best_ms = best_t;
    if (n_tried) *n_tried = tried;
    return LDPC_OK;
}

int ldpc_pool_config(const ldpc_pool *p, int *words_per_block, int *threads, int *max_workgroups, int *uniform_deg,
                     int *batch, int *pairs) {
    if (!p) {
        ldpc::set_last_error("null pool");
        return LDPC_ERR_ARG;
    }
    if (words_per_block) *words_per_block = p->P;
    if (threads) *threads = p->threads;
    if (max_workgroups) *max_workgroups = p->max_grid;
    if (uniform_deg) *uniform_deg = p->uniform_deg;
    if (batch) *batch = p->batch;
    if (pairs) *pairs = p->pairs;
    return LDPC_OK;
}

void *ldpc_pool_stream(ldpc_pool *p) { return p ? (void *)p->stream : nullptr; }

void ldpc_pool_destroy(ldpc_pool *p) {
    if (p) (void)hipStreamSynchronize(p->stream);
    pool_free(p);
}

}  // extern "C"
