// api.hip -- the extern "C" boundary (include/ldpc_mi355x.h) over the HIP
// kernels: decoder objects, the state API, device layout helpers, timer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ldpc_mi355x.h"
#include "code.h"
#include "kernels.h"
#include "pool.h"

using ldpc::Plan;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return fail(LDPC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

bool device_ok(std::string *why) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        *why = "no HIP device visible";
        return false;
    }
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        *why = "cannot query the HIP device";
        return false;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        *why = std::string("device is ") + prop.gcnArchName + ", library is built for gfx950 only";
        return false;
    }
    return true;
}

template <typename T>
hipError_t dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void **)p, count * sizeof(T));
}

}  // namespace

namespace ldpc {
void set_last_error(const std::string &m) { g_err = m; }
}  // namespace ldpc

struct ldpc_decoder {
    ldpc_code code;
    Plan plan;
    size_t W = 0;
    int variant = 0;
    int P = 1, log2P = 0, threads = 256;
    int uniform_deg = 0, batch = 1;  // single-degree kernel when uniform_deg > 0
    bool post_in_lds = true;
    // the word block does not fit LDS: fixed-iteration decodes without the
    // per-word stop run on a pair-mode frame queue (long_pool) instead
    bool long_frames = false;
    ldpc_pool *long_pool = nullptr;
    int8_t *d_frames = nullptr;   // OMS_v2 input back in frame-major order
    std::vector<int8_t> h_table;  // the learned table, for long_pool
    int early_term = 0;
    hipStream_t stream = nullptr;
    // host-buffer decode in chunks: copy-in, second compute and copy-out
    // streams, and the events that order the chunks (created on first use)
    hipStream_t s_in = nullptr, s_c2 = nullptr, s_out = nullptr;
    hipEvent_t ev_start = nullptr, ev_in[8] = {}, ev_dec[8] = {};
    // device
    uint2 *d_checks = nullptr;
    uint32_t *d_edge_off16 = nullptr;
    uint32_t *d_edge_var = nullptr;
    int *d_level_beg = nullptr;
    uint32_t *d_post = nullptr;
    uint32_t *d_msgs = nullptr;
    int8_t *d_in = nullptr;
    uint8_t *d_out = nullptr;
    uint8_t *d_stopped = nullptr;
    int *d_iters = nullptr;
    int8_t *d_table = nullptr;
    int tits = 0, tdeg = 0;
    // state API
    bool loaded = false;
    bool pending_first = false;
    int iters_done = 0;
};

struct ldpc_timer {
    hipStream_t stream = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
};

namespace {

void free_decoder(ldpc_decoder *d) {
    if (!d) return;
    (void)hipFree(d->d_checks);
    (void)hipFree(d->d_edge_var);
    (void)hipFree(d->d_level_beg);
    (void)hipFree(d->d_post);
    (void)hipFree(d->d_msgs);
    (void)hipFree(d->d_in);
    (void)hipFree(d->d_out);
    (void)hipFree(d->d_stopped);
    (void)hipFree(d->d_iters);
    (void)hipFree(d->d_edge_off16);
    (void)hipFree(d->d_table);
    (void)hipFree(d->d_frames);
    if (d->long_pool) ldpc_pool_destroy(d->long_pool);
    for (hipEvent_t e : d->ev_in)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : d->ev_dec)
        if (e) (void)hipEventDestroy(e);
    if (d->ev_start) (void)hipEventDestroy(d->ev_start);
    for (hipStream_t s : {d->s_in, d->s_c2, d->s_out})
        if (s) (void)hipStreamDestroy(s);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

__global__ void k_fill_u8(uint8_t *p, size_t n, uint8_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_i32(int *p, size_t n, int v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

ldpc::DecodeArgs base_args(const ldpc_decoder *d) {
    ldpc::DecodeArgs a;
    std::memset(&a, 0, sizeof a);
    a.checks = d->d_checks;
    a.edge_var = d->d_edge_var;
    a.edge_off16 = d->d_edge_off16;
    a.level_beg = d->d_level_beg;
    a.n_levels = d->plan.n_levels;
    a.N = d->code.N;
    a.C = d->code.C;
    a.E = d->code.E;
    a.P = d->P;
    a.log2P = d->log2P;
    a.W = d->W;
    a.post_state = d->d_post;
    a.msgs = d->d_msgs;
    a.variant = d->variant;
    a.early_term = d->early_term;
    a.stopped = d->d_stopped;
    a.iters_run = d->d_iters;
    if (d->variant == LDPC_LOMS || d->variant == LDPC_LNMS) {
        a.ltable = d->d_table;
        a.tits = d->tits;
        a.tdeg = d->tdeg;
    }
    return a;
}

int launch(ldpc_decoder *d, const ldpc::DecodeArgs &a, hipStream_t s) {
    if ((d->variant == LDPC_LOMS || d->variant == LDPC_LNMS) && !d->d_table)
        return fail(LDPC_ERR_STATE, "learned variant needs ldpc_decoder_set_table first");
    HIP_TRY(ldpc::launch_decode(a, d->plan.max_deg, d->threads, d->post_in_lds, s, d->uniform_deg, d->batch));
    return LDPC_OK;
}

int reset_counters(ldpc_decoder *d, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u8, dim3(256), dim3(256), 0, s, d->d_stopped, d->W, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_i32, dim3(256), dim3(256), 0, s, d->d_iters, d->W, 0);
    HIP_TRY(hipGetLastError());
    return LDPC_OK;
}

int check_iters(int iterations) {
    if (iterations < 1) return fail(LDPC_ERR_ARG, "iterations must be >= 1");
    return LDPC_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

const char *ldpc_last_error(void) { return g_err.c_str(); }
const char *ldpc_version(void) { return "ldpc_mi355x 0.1 (gfx950)"; }

// ---- codes ----------------------------------------------------------------
ldpc_code *ldpc_code_create(int n_vars, int n_checks, const int *deg, const uint32_t *pos, int q1_split) {
    if (!deg || !pos || n_vars <= 0 || n_checks <= 0) {
        fail(LDPC_ERR_ARG, "ldpc_code_create: bad arguments");
        return nullptr;
    }
    ldpc_code *c = new ldpc_code();
    c->N = n_vars;
    c->C = n_checks;
    c->deg.assign(deg, deg + n_checks);
    size_t E = 0;
    for (int i = 0; i < n_checks; i++) E += deg[i] > 0 ? (size_t)deg[i] : 0;
    c->pos.assign(pos, pos + E);
    std::string err = ldpc::finish_code(*c, q1_split);
    if (!err.empty()) {
        delete c;
        fail(LDPC_ERR_ARG, err);
        return nullptr;
    }
    return c;
}

ldpc_code *ldpc_code_create_grouped(int n_vars, int deg1, int count1, int deg2, int count2, const uint32_t *pos) {
    if (deg1 < 1 || count1 < 1 || count2 < 0 || (count2 > 0 && deg2 < 1)) {
        fail(LDPC_ERR_ARG, "ldpc_code_create_grouped: bad degree groups");
        return nullptr;
    }
    std::vector<int> deg((size_t)count1, deg1);
    deg.insert(deg.end(), (size_t)count2, deg2);
    return ldpc_code_create(n_vars, count1 + count2, deg.data(), pos, count1);
}

static ldpc_code *wrap(std::string err, ldpc_code *c) {
    if (!err.empty()) {
        delete c;
        fail(LDPC_ERR_ARG, err);
        return nullptr;
    }
    return c;
}

ldpc_code *ldpc_code_load_table(const char *path, int n_vars) {
    if (!path) { fail(LDPC_ERR_ARG, "null path"); return nullptr; }
    ldpc_code *c = new ldpc_code();
    return wrap(ldpc::load_table(*c, path, n_vars), c);
}
ldpc_code *ldpc_code_regular(int n_vars, int dv, int dc, uint64_t seed) {
    ldpc_code *c = new ldpc_code();
    return wrap(ldpc::gen_regular(*c, n_vars, dv, dc, seed), c);
}
ldpc_code *ldpc_code_qc(int mb, int nb, int z, double p, uint64_t seed) {
    ldpc_code *c = new ldpc_code();
    return wrap(ldpc::gen_qc(*c, mb, nb, z, p, seed), c);
}
ldpc_code *ldpc_code_ira(int k_info, int n_hi, int d_hi, int d_lo, uint64_t seed) {
    ldpc_code *c = new ldpc_code();
    return wrap(ldpc::gen_ira(*c, k_info, n_hi, d_hi, d_lo, seed), c);
}
ldpc_code *ldpc_code_reorder_by_colour(const ldpc_code *code) {
    if (!code) { fail(LDPC_ERR_ARG, "null code"); return nullptr; }
    ldpc_code *c = new ldpc_code();
    return wrap(ldpc::reorder_by_colour(*code, *c), c);
}
void ldpc_code_destroy(ldpc_code *code) { delete code; }

int ldpc_code_info(const ldpc_code *code, int *n_vars, int *n_checks, int *n_edges, int *max_deg,
                   int *n_levels, int *q1_split) {
    if (!code) return fail(LDPC_ERR_ARG, "null code");
    if (n_vars) *n_vars = code->N;
    if (n_checks) *n_checks = code->C;
    if (n_edges) *n_edges = code->E;
    if (max_deg) *max_deg = code->max_deg;
    if (n_levels) *n_levels = ldpc::dependency_levels(*code, nullptr);
    if (q1_split) *q1_split = code->q1_split;
    return LDPC_OK;
}

int ldpc_code_levels(const ldpc_code *code, int granule, int *level_of) {
    if (!code) return fail(LDPC_ERR_ARG, "null code");
    const ldpc::Plan p = ldpc::make_plan(*code, granule);
    if (level_of)
        for (int l = 0; l < p.n_levels; l++)
            for (int s = p.level_beg[l]; s < p.level_beg[l + 1]; s++) level_of[p.order[s]] = l;
    return p.n_levels;
}

int ldpc_code_export(const ldpc_code *code, int *deg, uint32_t *pos) {
    if (!code) return fail(LDPC_ERR_ARG, "null code");
    if (deg) std::copy(code->deg.begin(), code->deg.end(), deg);
    if (pos) std::copy(code->pos.begin(), code->pos.end(), pos);
    return LDPC_OK;
}

// ---- decoders ---------------------------------------------------------------
ldpc_decoder *ldpc_decoder_create(const ldpc_code *code, size_t nb_frames_words, int variant,
                                  const ldpc_tuning *tuning) {
    std::string why;
    if (!code || nb_frames_words == 0 || variant < LDPC_MS || variant > LDPC_LNMS) {
        fail(LDPC_ERR_ARG, "ldpc_decoder_create: bad arguments");
        return nullptr;
    }
    if (code->max_deg > ldpc::MAX_DEG) {
        fail(LDPC_ERR_ARG, "check degree above " + std::to_string(ldpc::MAX_DEG) + " is not supported");
        return nullptr;
    }
    if (!device_ok(&why)) {
        fail(LDPC_ERR_NODEV, why);
        return nullptr;
    }
    ldpc_tuning tu = {0, 0, -1, -1, -1, 0};
    if (tuning) tu = *tuning;
    ldpc_decoder *d = new ldpc_decoder();
    d->code = *code;
    d->W = nb_frames_words;
    d->variant = variant;
    d->early_term = tu.early_term >= 0 ? (tu.early_term != 0) : (variant == LDPC_2NMS);
    const size_t post_word_bytes = (size_t)code->N * 4;
    const size_t lds_budget = ldpc::MAX_LDS_BYTES - 1024;
    d->post_in_lds = tu.post_in_lds >= 0 ? (tu.post_in_lds != 0) : (post_word_bytes <= lds_budget);
    d->long_frames = tu.post_in_lds < 0 && post_word_bytes > lds_budget;
    if (d->post_in_lds && post_word_bytes > lds_budget) {
        free_decoder(d);
        fail(LDPC_ERR_ARG, "posterior of one word does not fit in LDS");
        return nullptr;
    }
    int P = tu.words_per_block;
    if (P <= 0) {
        P = 1;
        const size_t target = d->post_in_lds ? 40 * 1024 : (size_t)1 << 30;  // ~4 workgroups per CU
        while (P * 2 <= ldpc::MAX_P && post_word_bytes * (size_t)(P * 2) <= target && nb_frames_words % (size_t)(P * 2) == 0)
            P *= 2;
    }
    if (P < 1 || P > ldpc::MAX_P || (P & (P - 1)) || nb_frames_words % (size_t)P ||
        (d->post_in_lds && post_word_bytes * (size_t)P > lds_budget)) {
        free_decoder(d);
        fail(LDPC_ERR_ARG, "words_per_block must be a power of two <= 16 dividing nb_frames and fitting LDS");
        return nullptr;
    }
    d->P = P;
    while ((1 << d->log2P) < P) d->log2P++;
    // a level's (check, word) items fill 64-lane waves: sizes in multiples of 64 / P checks
    d->plan = ldpc::make_plan(d->code, 64 / P);
    int T = tu.threads > 0 ? tu.threads : 256;
    if (T % 64 || T % P || T < 64 || T > 1024) {
        free_decoder(d);
        fail(LDPC_ERR_ARG, "threads must be a multiple of 64 and of words_per_block, at most 1024");
        return nullptr;
    }
    d->threads = T;

    // single-degree kernel: every check of degree D, LDS offsets fit 16 bits,
    // and no check is subject to the OMS first-iteration no-clamp group
    int dmin = ldpc::MAX_DEG + 1;
    for (int x : code->deg) dmin = std::min(dmin, x);
    const bool uni_ok = dmin == code->max_deg && ldpc::uniform_degree_supported(dmin) && d->post_in_lds &&
                        (size_t)code->N * (size_t)P * 4 <= 65535 && (variant != LDPC_OMS || code->q1_split >= code->C);
    if (tu.uniform == 1 && !uni_ok) {
        free_decoder(d);
        fail(LDPC_ERR_ARG, "single-degree kernel requested but the code/blocking does not qualify");
        return nullptr;
    }
    d->uniform_deg = (tu.uniform != 0 && uni_ok) ? dmin : 0;
    if (T > ldpc::max_threads_for_degree(code->max_deg, d->uniform_deg > 0)) {
        if (tu.threads > 0) {
            free_decoder(d);
            fail(LDPC_ERR_ARG, "threads above the kernel's limit for this code (" +
                                   std::to_string(ldpc::max_threads_for_degree(code->max_deg, d->uniform_deg > 0)) + ")");
            return nullptr;
        }
        T = ldpc::max_threads_for_degree(code->max_deg, d->uniform_deg > 0);
    }
    d->threads = T;
    d->batch = tu.batch > 0 ? tu.batch : 2;
    if (d->batch != 1 && d->batch != 2 && d->batch != 4) {
        free_decoder(d);
        fail(LDPC_ERR_ARG, "batch must be 1, 2 or 4");
        return nullptr;
    }
    std::vector<uint32_t> off16;
    if (d->uniform_deg) {
        const int D = d->uniform_deg, H = (D + 1) / 2;
        off16.assign((size_t)code->C * H, 0);
        for (int cs = 0; cs < code->C; cs++)
            for (int j = 0; j < D; j++) {
                const uint32_t o = (d->plan.edge_var[(size_t)cs * D + j] << d->log2P) * 4;  // LDS byte offset
                off16[(size_t)cs * H + j / 2] |= (j & 1) ? (o << 16) : o;
            }
    }

    const size_t W = d->W;
    const ldpc_code &c = d->code;
    std::vector<uint2> ck((size_t)c.C);
    for (int i = 0; i < c.C; i++) ck[i] = make_uint2(d->plan.check_info[i], d->plan.check_e0[i]);
    hipError_t e = hipSuccess;
#define ALLOC(ptr, n) if (e == hipSuccess) e = dalloc(&d->ptr, (n))
    ALLOC(d_checks, (size_t)c.C);
    ALLOC(d_edge_var, (size_t)c.E);
    ALLOC(d_level_beg, (size_t)d->plan.n_levels + 1);
    ALLOC(d_post, (size_t)c.N * W);
    ALLOC(d_msgs, (size_t)c.E * W);
    ALLOC(d_in, (size_t)c.N * W * 4);
    ALLOC(d_out, (size_t)c.N * W * 4);
    ALLOC(d_stopped, W);
    ALLOC(d_iters, W);
    ALLOC(d_edge_off16, off16.size());
#undef ALLOC
    // a blocking stream: work the caller enqueues on the null stream stays ordered with it
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&d->stream, hipStreamDefault);
    if (e == hipSuccess) e = hipMemcpy(d->d_checks, ck.data(), ck.size() * sizeof(uint2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d->d_edge_var, d->plan.edge_var.data(), (size_t)c.E * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && !off16.empty())
        e = hipMemcpy(d->d_edge_off16, off16.data(), off16.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d->d_level_beg, d->plan.level_beg.data(), d->plan.level_beg.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d->d_msgs, 0, (size_t)c.E * W * 4);
    if (e == hipSuccess) e = hipMemset(d->d_stopped, 0, W);
    if (e == hipSuccess) e = hipMemset(d->d_iters, 0, W * sizeof(int));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        const bool oom = e == hipErrorOutOfMemory;
        free_decoder(d);
        fail(oom ? LDPC_ERR_NOMEM : LDPC_ERR_HIP, std::string("ldpc_decoder_create: ") + hipGetErrorString(e));
        return nullptr;
    }
    return d;
}

int ldpc_decoder_initialize(ldpc_decoder *dec) {
    if (!dec) return fail(LDPC_ERR_ARG, "null decoder");
    return LDPC_OK;
}

void ldpc_decoder_destroy(ldpc_decoder *dec) {
    if (dec) (void)hipStreamSynchronize(dec->stream);
    free_decoder(dec);
}

void *ldpc_decoder_stream(ldpc_decoder *dec) { return dec ? (void *)dec->stream : nullptr; }

int ldpc_decoder_config(const ldpc_decoder *dec, int *words_per_block, int *threads, int *post_in_lds,
                        int *n_levels, int *uniform_deg, int *batch) {
    if (!dec) return fail(LDPC_ERR_ARG, "null decoder");
    if (words_per_block) *words_per_block = dec->P;
    if (threads) *threads = dec->threads;
    if (post_in_lds) *post_in_lds = dec->post_in_lds ? 1 : 0;
    if (n_levels) *n_levels = dec->plan.n_levels;
    if (uniform_deg) *uniform_deg = dec->uniform_deg;
    if (batch) *batch = dec->batch;
    return LDPC_OK;
}

int ldpc_decoder_set_table(ldpc_decoder *dec, const int8_t *table, int n_iters, int n_degs) {
    // the frame-queue decoder's limits (long frames run on one), checked
    // before any state changes
    if (!dec || !table || n_iters < 1 || n_degs < 1 || n_iters > 64 || n_degs > ldpc::MAX_DEG + 1)
        return fail(LDPC_ERR_ARG, "ldpc_decoder_set_table: bad table (at most 64 rows, 33 columns)");
    const int lo = 0, hi = dec->variant == LDPC_LNMS ? 8 : 127;
    for (int i = 0; i < n_iters * n_degs; i++)
        if (table[i] < lo || table[i] > hi) return fail(LDPC_ERR_ARG, "table value out of range");
    HIP_TRY(hipStreamSynchronize(dec->stream));
    (void)hipFree(dec->d_table);
    dec->d_table = nullptr;
    HIP_TRY(dalloc(&dec->d_table, (size_t)n_iters * n_degs));
    HIP_TRY(hipMemcpy(dec->d_table, table, (size_t)n_iters * n_degs, hipMemcpyHostToDevice));
    dec->tits = n_iters;
    dec->tdeg = n_degs;
    dec->h_table.assign(table, table + (size_t)n_iters * n_degs);
    if (dec->long_pool && ldpc_pool_set_table(dec->long_pool, table, n_iters, n_degs) != LDPC_OK) {
        // the pair block no longer fits beside this table: drop the frame
        // queue, long frames then use the global-memory-posterior kernel
        ldpc_pool_destroy(dec->long_pool);
        dec->long_pool = nullptr;
        dec->long_frames = false;
    }
    return LDPC_OK;
}

static size_t io_bytes(const ldpc_decoder *d) { return (size_t)d->code.N * d->W * 4; }

// Long frames (word block larger than LDS), fixed iterations, no per-word
// stop: each frame is decoded for exactly `iterations` iterations by the
// pair-mode frame queue (two frames per 16-bit lane, posterior in LDS) with
// no stop test -- the same per-frame layered decode, hence the same
// decisions as the global-memory-posterior kernel.  Not for OMS with two
// degree groups (its first-iteration rule is per word group, not per frame).
static bool use_long_pool(const ldpc_decoder *d) {
    return d->long_frames && !d->early_term && !(d->variant == LDPC_OMS && d->code.q1_split < d->code.C);
}
static int decode_long(ldpc_decoder *dec, const int8_t *d_in, uint8_t *d_out, int iterations, hipStream_t s) {
    if ((dec->variant == LDPC_LOMS || dec->variant == LDPC_LNMS) && dec->h_table.empty())
        return fail(LDPC_ERR_STATE, "learned variant needs ldpc_decoder_set_table first");
    if (!dec->long_pool) {
        const ldpc_tuning tu = {0, 0, -1, 2, 0, 0};  // pairs, automatic block and workgroup size
        dec->long_pool = ldpc_pool_create(&dec->code, dec->variant, iterations, LDPC_STOP_NONE, &tu);
        if (!dec->long_pool) return LDPC_ERR_ARG;  // (the caller falls back to the global-memory kernel)
        if (!dec->h_table.empty() &&
            ldpc_pool_set_table(dec->long_pool, dec->h_table.data(), dec->tits, dec->tdeg) != LDPC_OK) {
            ldpc_pool_destroy(dec->long_pool);  // no room beside the table: the global-memory kernel
            dec->long_pool = nullptr;
            return LDPC_ERR_ARG;
        }
    }
    dec->long_pool->max_iters = iterations;
    const int8_t *frames = d_in;
    if (dec->variant == LDPC_OMS_V2) {  // input in the device word layout
        if (!dec->d_frames) HIP_TRY(dalloc(&dec->d_frames, io_bytes(dec)));
        HIP_TRY(ldpc::launch_deinterleave_raw(reinterpret_cast<const uint32_t *>(d_in), dec->d_frames, dec->code.N,
                                              (int)dec->W, s));
        frames = dec->d_frames;
    }
    return ldpc_pool_decode(dec->long_pool, frames, 4 * (uint64_t)dec->W, d_out, nullptr, 0, s);
}

int ldpc_decoder_decode_device(ldpc_decoder *dec, const int8_t *d_in, uint8_t *d_out, int iterations, void *stream) {
    if (!dec || !d_in || !d_out) return fail(LDPC_ERR_ARG, "null argument");
    int rc = check_iters(iterations);
    if (rc) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : dec->stream;
    if (use_long_pool(dec)) {
        dec->loaded = false;
        rc = decode_long(dec, d_in, d_out, iterations, s);
        if (rc != LDPC_ERR_ARG || dec->long_pool) return rc;
        dec->long_frames = false;  // no pair block fits either: global-memory posterior
    }
    if (dec->early_term) {
        rc = reset_counters(dec, s);
        if (rc) return rc;
    }
    ldpc::DecodeArgs a = base_args(dec);
    if (dec->variant == LDPC_OMS_V2) {
        a.load_mode = ldpc::LOAD_WORDS;
        a.in_words = reinterpret_cast<const uint32_t *>(d_in);
    } else {
        a.load_mode = ldpc::LOAD_FRAMES;
        a.in_frames = d_in;
    }
    a.out_hard = d_out;
    a.store_state = 0;
    a.iters = iterations;
    a.first = 1;
    a.iter_base = 0;
    if (!dec->early_term) { a.stopped = nullptr; a.iters_run = nullptr; }
    dec->loaded = false;  // the state buffers no longer describe a loaded decode
    return launch(dec, a, s);
}

// Host-buffer decode of a large batch in K chunks of word blocks: chunk k's
// H2D (copy-in stream), its decode (the decoder's stream and a second compute
// stream in turn, so that one chunk's kernel tail overlaps the next one's
// start) and its D2H (copy-out stream) are ordered by events, so the copies
// of one chunk run under the decode of another.  Each chunk's frames are one
// contiguous byte range of the frame-major buffers (OMS_v2's word-layout
// input: a column range of every variable's row, one 2-D copy).
constexpr int DECODE_CHUNKS = 8;
static int chunk_streams(ldpc_decoder *dec) {  // creates what is missing (idempotent)
    for (hipStream_t *s : {&dec->s_in, &dec->s_c2, &dec->s_out})
        if (!*s) HIP_TRY(hipStreamCreateWithFlags(s, hipStreamDefault));
    for (hipEvent_t *e : {&dec->ev_start})
        if (!*e) HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    for (int k = 0; k < DECODE_CHUNKS; k++) {
        if (!dec->ev_in[k]) HIP_TRY(hipEventCreateWithFlags(&dec->ev_in[k], hipEventDisableTiming));
        if (!dec->ev_dec[k]) HIP_TRY(hipEventCreateWithFlags(&dec->ev_dec[k], hipEventDisableTiming));
    }
    return LDPC_OK;
}
static int enqueue_chunks(ldpc_decoder *dec, const int8_t *in, uint8_t *out, int iterations) {
    if (dec->early_term) {
        const int rc = reset_counters(dec, dec->stream);
        if (rc) return rc;
    }
    // everything enqueued on the decoder's stream before this call comes first
    HIP_TRY(hipEventRecord(dec->ev_start, dec->stream));
    for (hipStream_t s : {dec->s_in, dec->s_c2, dec->s_out}) HIP_TRY(hipStreamWaitEvent(s, dec->ev_start, 0));
    const size_t N = (size_t)dec->code.N, P = (size_t)dec->P, nblk = dec->W / P;
    const bool words = dec->variant == LDPC_OMS_V2;
    for (int k = 0; k < DECODE_CHUNKS; k++) {
        const size_t b0 = nblk * k / DECODE_CHUNKS, b1 = nblk * (k + 1) / DECODE_CHUNKS;
        const size_t f0 = b0 * P * 4 * N, fb = (b1 - b0) * P * 4 * N;  // frame-major byte range
        if (words)
            HIP_TRY(hipMemcpy2DAsync(dec->d_in + b0 * P * 4, dec->W * 4, in + b0 * P * 4, dec->W * 4, (b1 - b0) * P * 4,
                                     N, hipMemcpyHostToDevice, dec->s_in));
        else
            HIP_TRY(hipMemcpyAsync(dec->d_in + f0, in + f0, fb, hipMemcpyHostToDevice, dec->s_in));
        HIP_TRY(hipEventRecord(dec->ev_in[k], dec->s_in));
        hipStream_t cs = (k & 1) ? dec->s_c2 : dec->stream;
        HIP_TRY(hipStreamWaitEvent(cs, dec->ev_in[k], 0));
        ldpc::DecodeArgs a = base_args(dec);
        if (words) {
            a.load_mode = ldpc::LOAD_WORDS;
            a.in_words = reinterpret_cast<const uint32_t *>(dec->d_in);
        } else {
            a.load_mode = ldpc::LOAD_FRAMES;
            a.in_frames = dec->d_in;
        }
        a.out_hard = dec->d_out;
        a.iters = iterations;
        a.first = 1;
        if (!dec->early_term) { a.stopped = nullptr; a.iters_run = nullptr; }
        a.block0 = b0;
        a.nblocks = (unsigned)(b1 - b0);
        const int rc = launch(dec, a, cs);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(dec->ev_dec[k], cs));
        HIP_TRY(hipStreamWaitEvent(dec->s_out, dec->ev_dec[k], 0));
        HIP_TRY(hipMemcpyAsync(out + f0, dec->d_out + f0, fb, hipMemcpyDeviceToHost, dec->s_out));
    }
    return LDPC_OK;
}
static int decode_chunked(ldpc_decoder *dec, const int8_t *in, uint8_t *out, int iterations) {
    int rc = chunk_streams(dec);
    if (rc) return rc;
    rc = enqueue_chunks(dec, in, out, iterations);
    // the copy-out stream waited on every chunk's decode, which waited on its
    // copy-in; after an error, whatever was enqueued still ends before the
    // call returns (the caller's buffers may be freed right after)
    if (rc == LDPC_OK) {
        HIP_TRY(hipStreamSynchronize(dec->s_out));
    } else {
        for (hipStream_t s : {dec->s_in, dec->stream, dec->s_c2, dec->s_out}) (void)hipStreamSynchronize(s);
    }
    dec->loaded = false;
    return rc;
}

int ldpc_decoder_decode(ldpc_decoder *dec, const int8_t *in, uint8_t *out, int iterations) {
    if (!dec || !in || !out) return fail(LDPC_ERR_ARG, "null argument");
    int rc = check_iters(iterations);
    if (rc) return rc;
    // chunks of at least 4 MiB and 512 workgroups each (the two compute
    // streams fill each other's kernel tails)
    if (!use_long_pool(dec) && io_bytes(dec) >= ((size_t)DECODE_CHUNKS << 22) &&
        dec->W / dec->P >= (size_t)DECODE_CHUNKS * 512)
        return decode_chunked(dec, in, out, iterations);
    HIP_TRY(hipMemcpyAsync(dec->d_in, in, io_bytes(dec), hipMemcpyHostToDevice, dec->stream));
    rc = ldpc_decoder_decode_device(dec, dec->d_in, dec->d_out, iterations, dec->stream);
    if (rc) {
        (void)hipStreamSynchronize(dec->stream);  // the copy-in reads `in`: it ends before the call returns
        return rc;
    }
    HIP_TRY(hipMemcpyAsync(out, dec->d_out, io_bytes(dec), hipMemcpyDeviceToHost, dec->stream));
    HIP_TRY(hipStreamSynchronize(dec->stream));
    return LDPC_OK;
}

int ldpc_decoder_load(ldpc_decoder *dec, const int8_t *in) {
    if (!dec || !in) return fail(LDPC_ERR_ARG, "null argument");
    HIP_TRY(hipMemcpyAsync(dec->d_in, in, io_bytes(dec), hipMemcpyHostToDevice, dec->stream));
    int rc = reset_counters(dec, dec->stream);
    if (rc) return rc;
    ldpc::DecodeArgs a = base_args(dec);
    if (dec->variant == LDPC_OMS_V2) {
        a.load_mode = ldpc::LOAD_WORDS;
        a.in_words = reinterpret_cast<const uint32_t *>(dec->d_in);
    } else {
        a.load_mode = ldpc::LOAD_FRAMES;
        a.in_frames = dec->d_in;
    }
    a.store_state = 1;
    a.iters = 0;
    a.first = 1;
    a.iters_run = nullptr;
    rc = launch(dec, a, dec->stream);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(dec->stream));
    dec->loaded = true;
    dec->pending_first = true;
    dec->iters_done = 0;
    return LDPC_OK;
}

int ldpc_decoder_iterate(ldpc_decoder *dec, int iterations) {
    if (!dec) return fail(LDPC_ERR_ARG, "null decoder");
    int rc = check_iters(iterations);
    if (rc) return rc;
    if (!dec->loaded) return fail(LDPC_ERR_STATE, "ldpc_decoder_iterate before ldpc_decoder_load");
    ldpc::DecodeArgs a = base_args(dec);
    a.load_mode = ldpc::LOAD_STATE;
    a.store_state = 1;
    a.iters = iterations;
    a.first = dec->pending_first ? 1 : 0;
    a.iter_base = dec->iters_done;
    rc = launch(dec, a, dec->stream);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(dec->stream));
    dec->pending_first = false;
    dec->iters_done += iterations;
    return LDPC_OK;
}

int ldpc_decoder_get_state(ldpc_decoder *dec, uint32_t *post, uint32_t *msg) {
    if (!dec) return fail(LDPC_ERR_ARG, "null decoder");
    if (!dec->loaded) return fail(LDPC_ERR_STATE, "no loaded decode");
    const ldpc_code &c = dec->code;
    const size_t W = dec->W, P = (size_t)dec->P, nb = W / P;
    HIP_TRY(hipStreamSynchronize(dec->stream));
    if (post) {
        std::vector<uint32_t> h((size_t)c.N * W);
        HIP_TRY(hipMemcpy(h.data(), dec->d_post, h.size() * 4, hipMemcpyDeviceToHost));
        for (size_t b = 0; b < nb; b++)
            for (int v = 0; v < c.N; v++)
                for (size_t w = 0; w < P; w++) post[(size_t)v * W + b * P + w] = h[(b * c.N + v) * P + w];
    }
    if (msg) {
        std::vector<uint32_t> h((size_t)c.E * W);
        HIP_TRY(hipMemcpy(h.data(), dec->d_msgs, h.size() * 4, hipMemcpyDeviceToHost));
        for (int t = 0; t < c.C; t++) {
            const int s = dec->plan.slot_of[t];
            const size_t e0 = dec->plan.check_e0[s];
            const size_t D = (size_t)c.deg[t];
            for (size_t j = 0; j < D; j++) {
                const size_t e = c.eoff[t] + j;
                for (size_t b = 0; b < nb; b++)
                    for (size_t w = 0; w < P; w++)
                        msg[e * W + b * P + w] = h[b * (size_t)c.E * P + e0 * P + w * D + j];
            }
        }
    }
    return LDPC_OK;
}

int ldpc_decoder_get_hard(ldpc_decoder *dec, uint8_t *out) {
    if (!dec || !out) return fail(LDPC_ERR_ARG, "null argument");
    if (!dec->loaded) return fail(LDPC_ERR_STATE, "no loaded decode");
    ldpc::DecodeArgs a = base_args(dec);
    a.load_mode = ldpc::LOAD_STATE;
    a.store_state = 0;
    a.iters = 0;
    a.out_hard = dec->d_out;
    a.iters_run = nullptr;
    int rc = launch(dec, a, dec->stream);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, dec->d_out, io_bytes(dec), hipMemcpyDeviceToHost, dec->stream));
    HIP_TRY(hipStreamSynchronize(dec->stream));
    return LDPC_OK;
}

int ldpc_decoder_get_stop_info(ldpc_decoder *dec, uint8_t *stopped, int *iters_run) {
    if (!dec) return fail(LDPC_ERR_ARG, "null decoder");
    HIP_TRY(hipStreamSynchronize(dec->stream));
    if (stopped) HIP_TRY(hipMemcpy(stopped, dec->d_stopped, dec->W, hipMemcpyDeviceToHost));
    if (iters_run) HIP_TRY(hipMemcpy(iters_run, dec->d_iters, dec->W * sizeof(int), hipMemcpyDeviceToHost));
    return LDPC_OK;
}

// ---- device layout helpers -------------------------------------------------
int ldpc_interleave_device(const int8_t *d_in, uint32_t *d_words, int n_vars, size_t W, void *stream) {
    if (!d_in || !d_words || n_vars <= 0 || W == 0 || W > 0x7FFFFFFF) return fail(LDPC_ERR_ARG, "bad arguments");
    HIP_TRY(ldpc::launch_interleave(d_in, d_words, n_vars, (int)W, (hipStream_t)stream));
    return LDPC_OK;
}
int ldpc_deinterleave_hd_device(const uint32_t *d_words, uint8_t *d_out, int n_vars, size_t W, void *stream) {
    if (!d_words || !d_out || n_vars <= 0 || W == 0 || W > 0x7FFFFFFF) return fail(LDPC_ERR_ARG, "bad arguments");
    HIP_TRY(ldpc::launch_deinterleave_hd(d_words, d_out, n_vars, (int)W, (hipStream_t)stream));
    return LDPC_OK;
}

// ---- channel and error counting ------------------------------------------------
int ldpc_channel_generate(void *d_out, int layout, int n_vars, size_t frames, uint64_t first_frame, uint64_t seed,
                          uint32_t snr_idx, double sigma, int quant, void *stream) {
    if (!d_out || n_vars <= 0 || frames == 0 || (layout != 0 && layout != 1) || (quant != 0 && quant != 1) ||
        !(sigma > 0.0))
        return fail(LDPC_ERR_ARG, "ldpc_channel_generate: bad arguments");
    if (layout == 1 && frames % 4) return fail(LDPC_ERR_ARG, "device-word layout needs frames % 4 == 0");
    HIP_TRY(ldpc::launch_channel(d_out, layout, n_vars, frames, first_frame, seed, snr_idx, sigma, quant,
                                 (hipStream_t)stream));
    return LDPC_OK;
}

double ldpc_channel_sigma(double snr_db, double rate, int es_n0) {
    double ebn0 = es_n0 ? snr_db - 10.0 * std::log10(2.0 * rate) : snr_db;
    const double interm = -0.1 * (ebn0 + 10.0 * std::log10(rate));
    return std::sqrt(std::pow(10.0, interm) / 2.0);
}

int ldpc_decoder_count_errors(ldpc_decoder *dec, const uint8_t *d_hard, size_t frames, int n_info,
                              uint64_t *d_counters, uint32_t *d_frame_bit_errors, uint8_t *d_frame_syndrome_ok,
                              int accumulate, void *stream) {
    if (!dec || !d_hard || !d_counters || frames == 0 || n_info < 0 || n_info > dec->code.N)
        return fail(LDPC_ERR_ARG, "ldpc_decoder_count_errors: bad arguments");
    if ((size_t)dec->code.N > ldpc::MAX_LDS_BYTES - 1024)
        return fail(LDPC_ERR_ARG, "frame longer than the LDS staging buffer");
    HIP_TRY(ldpc::set_count_lds_limit());  // once per process
    HIP_TRY(ldpc::launch_count(d_hard, dec->code.N, frames, n_info, dec->d_checks, dec->d_edge_var, dec->code.C,
                               reinterpret_cast<unsigned long long *>(d_counters), d_frame_bit_errors,
                               d_frame_syndrome_ok, accumulate == 0, (hipStream_t)stream));
    return LDPC_OK;
}

// ---- simulation loop (src/main.cpp:237-335 for one Eb/N0 point) -----------------
int ldpc_simulate(ldpc_decoder *dec, double ebn0_db, int quant, uint64_t seed, uint32_t snr_idx, uint64_t n_frames,
                  int iterations, ldpc_sim_stats *out) {
    if (!dec || !out || n_frames == 0) return fail(LDPC_ERR_ARG, "ldpc_simulate: bad arguments");
    int rc = check_iters(iterations);
    if (rc) return rc;
    const ldpc_code &c = dec->code;
    const double rate = (double)(c.N - c.C) / (double)c.N;
    const double sigma = ldpc_channel_sigma(ebn0_db, rate, 0);
    const size_t batch = 4 * dec->W;
    const uint64_t batches = (n_frames + batch - 1) / batch;
    uint64_t *d_cnt = nullptr;
    HIP_TRY(hipMalloc((void **)&d_cnt, 3 * sizeof(uint64_t)));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    std::vector<int> iters(dec->W);
    double ms = 0.0, it_sum = 0.0;
    hipStream_t s = dec->stream;
    for (uint64_t k = 0; k < batches && rc == LDPC_OK; k++) {
        rc = ldpc_channel_generate(dec->d_in, dec->variant == LDPC_OMS_V2 ? 1 : 0, c.N, batch, k * batch, seed,
                                   snr_idx, sigma, quant, s);
        if (rc) break;
        (void)hipEventRecord(e0, s);
        rc = ldpc_decoder_decode_device(dec, dec->d_in, dec->d_out, iterations, s);
        (void)hipEventRecord(e1, s);
        if (rc) break;
        rc = ldpc_decoder_count_errors(dec, dec->d_out, batch, c.N - c.C, d_cnt, nullptr, nullptr, k > 0, s);
        if (rc) break;
        if (hipEventSynchronize(e1) != hipSuccess) { rc = fail(LDPC_ERR_HIP, "event sync"); break; }
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        ms += t;
        if (dec->early_term) {
            if (hipMemcpyAsync(iters.data(), dec->d_iters, dec->W * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) { rc = fail(LDPC_ERR_HIP, "iteration counts"); break; }
            for (int x : iters) it_sum += 4.0 * x;
        } else {
            it_sum += (double)batch * iterations;
        }
    }
    uint64_t cnt[3] = {0, 0, 0};
    if (rc == LDPC_OK && hipMemcpy(cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(LDPC_ERR_HIP, "counter copy");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(d_cnt);
    if (rc) return rc;
    out->frames = batches * batch;
    out->bit_errors = cnt[0];
    out->frame_errors = cnt[1];
    out->syndrome_failures = cnt[2];
    out->avg_iterations = it_sum / (double)out->frames;
    out->decode_ms = ms;
    return LDPC_OK;
}

// ---- memory helpers -----------------------------------------------------------
void *ldpc_device_alloc(size_t bytes) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        fail(e == hipErrorOutOfMemory ? LDPC_ERR_NOMEM : LDPC_ERR_HIP, hipGetErrorString(e));
        return nullptr;
    }
    return p;
}
void ldpc_device_free(void *p) { (void)hipFree(p); }
void *ldpc_host_alloc_pinned(size_t bytes) {
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        fail(LDPC_ERR_HIP, hipGetErrorString(e));
        return nullptr;
    }
    return p;
}
void ldpc_host_free_pinned(void *p) { (void)hipHostFree(p); }
int ldpc_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return LDPC_OK;
}
int ldpc_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return LDPC_OK;
}
int ldpc_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return LDPC_OK;
}
int ldpc_device_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return LDPC_OK;
}
int ldpc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    int good = 0;
    for (int i = 0; i < n; i++) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, i) == hipSuccess && std::strncmp(p.gcnArchName, "gfx950", 6) == 0) good++;
    }
    return good;
}

// ---- timer ------------------------------------------------------------------------
ldpc_timer *ldpc_timer_create(void *stream) {
    ldpc_timer *t = new ldpc_timer();
    t->stream = (hipStream_t)stream;
    if (hipEventCreate(&t->a) != hipSuccess || hipEventCreate(&t->b) != hipSuccess) {
        fail(LDPC_ERR_HIP, "hipEventCreate failed");
        if (t->a) (void)hipEventDestroy(t->a);
        delete t;
        return nullptr;
    }
    return t;
}
int ldpc_timer_start(ldpc_timer *t) {
    if (!t) return fail(LDPC_ERR_ARG, "null timer");
    HIP_TRY(hipEventRecord(t->a, t->stream));
    return LDPC_OK;
}
int ldpc_timer_stop(ldpc_timer *t) {
    if (!t) return fail(LDPC_ERR_ARG, "null timer");
    HIP_TRY(hipEventRecord(t->b, t->stream));
    return LDPC_OK;
}
double ldpc_timer_elapsed_ms(ldpc_timer *t) {
    if (!t) return -1.0;
    if (hipEventSynchronize(t->b) != hipSuccess) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t->a, t->b) != hipSuccess) return -1.0;
    return ms;
}
void ldpc_timer_destroy(ldpc_timer *t) {
    if (!t) return;
    (void)hipEventDestroy(t->a);
    (void)hipEventDestroy(t->b);
    delete t;
}

}  // extern "C"
