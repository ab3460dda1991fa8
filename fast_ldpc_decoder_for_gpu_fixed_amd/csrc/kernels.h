// kernels.h -- launch interface between the C-ABI host code and the HIP
// kernels of the decoding path.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>

namespace ldpc {

// Runs f once per process (thread-safe) and returns its result on every call;
// used for the kernels' LDS attribute, which is set before their first launch.
template <typename F>
inline hipError_t once_per_process(std::once_flag &flag, hipError_t &result, F f) {
    std::call_once(flag, [&] { result = f(); });
    return result;
}

// numbering shared with include/ldpc_mi355x.h
enum Variant { VAR_MS = 0, VAR_OMS = 1, VAR_OMS_V2 = 2, VAR_NMS = 3, VAR_2NMS = 4, VAR_LOMS = 5, VAR_LNMS = 6 };
enum LoadMode { LOAD_FRAMES = 0, LOAD_WORDS = 1, LOAD_STATE = 2 };

constexpr int MAX_P = 16;                  // words per workgroup block
constexpr int MAX_DEG = 32;                // largest check degree handled
// The frame-queue kernels gather a check's variable ids and messages for
// the largest degree of their class without a bound test, so the edge-id and
// message arrays carry this many spare entries (u32) past their end.
constexpr int GATHER_PAD = MAX_DEG;
constexpr size_t MAX_LDS_BYTES = 160 * 1024;

struct DecodeArgs {
    // code plan (level-sorted)
    const uint2 *checks;       // [C] x = degree | q1 << 8, y = first edge (level order)
    const uint32_t *edge_var;  // [E] variable index per level-order edge
    const uint32_t *edge_off16;  // uniform-degree codes: per check (D+1)/2 dwords of
                                 // LDS byte offsets v*P*4 packed two per dword, or null
    const int *level_beg;      // [n_levels + 1]
    int n_levels, N, C, E;
    // blocking
    int P, log2P;              // words per workgroup (power of two)
    size_t W;                  // words in the batch (multiple of P)
    size_t block0;             // the launch covers word blocks block0 ..
    unsigned nblocks;          //   block0 + nblocks - 1 (nblocks 0: all W/P)
    // state
    uint32_t *post_state;      // [W/P][N][P] words (internal posterior)
    uint32_t *msgs;            // [W/P][E][P] words, per check [word][j]
    // input / output
    const int8_t *in_frames;   // frame-major LLRs, 4W*N bytes (LOAD_FRAMES)
    const uint32_t *in_words;  // var-major words [N][W]       (LOAD_WORDS)
    uint8_t *out_hard;         // frame-major decisions, or null
    int load_mode;
    int store_state;           // write the posterior back to post_state
    // iteration control
    int iters, first, variant, early_term;
    uint8_t *stopped;          // [W] per-word stop flags or null
    int *iters_run;            // [W] iterations run in this launch, or null
    const int8_t *ltable;      // learned table [tits][tdeg] or null
    int tits, tdeg, iter_base;
};

// uniform_deg > 0 selects the single-degree kernel (needs edge_off16 and the
// posterior in LDS); batch = items whose loads are issued together (1, 2, 4).
hipError_t launch_decode(const DecodeArgs &a, int dmax, int threads, bool post_in_lds, hipStream_t s,
                         int uniform_deg, int batch);
bool uniform_degree_supported(int d);
int max_threads_for_degree(int dmax, bool uniform);
hipError_t launch_interleave(const int8_t *in, uint32_t *out, int N, int W, hipStream_t s);
hipError_t launch_deinterleave_hd(const uint32_t *in, uint8_t *out, int N, int W, hipStream_t s);
hipError_t launch_deinterleave_raw(const uint32_t *in, int8_t *out, int N, int W, hipStream_t s);

// persistent frame-queue decoder (pool_kernels.hip)
struct PoolArgs {
    const uint2 *checks;
    const uint32_t *edge_var;
    const uint32_t *edge_lds;  // [E] + GATHER_PAD: LDS byte address of each level-order edge's variable
                               // (variable * 4P for words, * 2P for pairs)
    const int *level_beg;
    int n_levels, N, C, E;
    int P, log2P;
    const int8_t *in_frames;      // LLRs: frame f at in_frames + off(f) (N bytes)
    uint8_t *out_hard;            // decisions: frame f at out_hard + off(f)
    int *out_iters;               // [frames] iterations per frame, or null
    unsigned long long *next;     // queue head (device counter, starts at 0)
    unsigned long long total;     // frames in the queue (when total_ptr is null)
    const unsigned *total_ptr;    // device count of queued frames, or null
    const unsigned *frame_list;   // queue position -> frame id, or null (identity)
    const unsigned long long *frame_off;  // frame id -> byte offset, or null (f * N)
    uint32_t *msgs;               // [grid][E][P] per-workgroup message words (pairs: u16 units)
    int max_iters, variant;
    const int8_t *ltable;         // learned table or null
    int tits, tdeg;
    const uint32_t *edge_off16;   // single-degree path: packed LDS byte offsets, or null
    int uniform_deg, batch;       // degree of that path, items in flight per thread (1, 2, 4)
    int stop_rule;                // 0: syndrome of the decisions, 1: v2c parity (reference rule), 2: none
    int pairs;                    // 1: posterior and messages hold 2 frames per 16 bits (P = pairs)
    const uint32_t *rows;         // pair mode, degrees <= 8: [C][8] LDS byte offsets per check
                                  // (variable * 2P; unused entries: the padding variable N), and
                                  // messages as [grid][C][P][8] u16 runs; else null
    const int *level_deg;         // [n_levels] degree shared by every check of the level, or 0
};
// largest posterior block of a frame-queue workgroup; the slot bookkeeping and
// the learned table (< 4.4 KB) follow it in the same dynamic LDS allocation
constexpr size_t POOL_LDS_DYNAMIC = MAX_LDS_BYTES - 8192;
size_t pool_lds_bytes(const PoolArgs &a);
hipError_t launch_pool(const PoolArgs &a, int dmax, int grid, int threads, hipStream_t s);
// sets the LDS limit of the instantiation `a` selects (call before any launch
// and outside graph capture) and returns how many of its workgroups a CU holds
// (-1: a pair-mode rows kernel with static LDS, which its addressing forbids)
int pool_prepare(const PoolArgs &a, int dmax, int threads);
hipError_t launch_zero_u32(unsigned *p, int n, hipStream_t s);
// queue order least reliable first: scratch of order_bytes(frames) bytes
// (4-byte aligned); *list_out receives the queue-position -> frame-id list in it
size_t order_bytes(unsigned long long frames);
hipError_t launch_order(const int8_t *in, int N, unsigned long long frames, void *scratch, hipStream_t s,
                        const unsigned **list_out);

// float shadow decoder (learned tables)
struct ShadowArgs {
    const uint2 *checks;
    const uint32_t *edge_var;
    const int *level_beg;
    int n_levels, N, C, E;
    int FB, lfb;            // frames per workgroup (power of two)
    size_t frames;          // multiple of FB
    const int8_t *llr;      // frame-major [frames][N]
    float *msg;             // [frames/FB][E][FB], per check [frame][j]; gradients in reverse
    float *rstore;          // [T][frames/FB][E*FB] pre-clamp v2c values
    float *pend;            // [T][frames/FB][N*FB] posterior after each iteration
    const float *params;    // [T][dcols]
    float *grad;            // [T][dcols], accumulated with atomics
    float *loss;            // one float, accumulated with atomics
    float *post_out;        // optional [frames][N] posterior after T iterations
    int T, dcols, mode;     // mode 0 offset, 1 scale
};
hipError_t launch_shadow(const ShadowArgs &a, bool backward, hipStream_t s);

hipError_t launch_channel(void *out, int layout, int N, size_t frames, uint64_t f0, uint64_t seed, uint32_t snr_idx,
                          double sigma, int quant, hipStream_t s);
// per-frame Eb/N0 uniform in [lo, hi] dB (frame-major output)
hipError_t launch_channel_range(int8_t *out, int N, size_t frames, uint64_t f0, uint64_t seed, uint32_t snr_idx,
                                double lo_db, double hi_db, double rate, int quant, hipStream_t s);
hipError_t set_count_lds_limit();
hipError_t launch_count(const uint8_t *hard, int N, size_t frames, int n_info, const uint2 *checks,
                        const uint32_t *edge_var, int C, unsigned long long *counters, uint32_t *per_frame_bits,
                        uint8_t *per_frame_ok, bool zero_first, hipStream_t s);

}  // namespace ldpc
