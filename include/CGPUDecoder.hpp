// CGPUDecoder.hpp -- the reference's decoder-object interface, header-only,
// over the C ABI of libldpc_mi355x.so (ldpc_mi355x.h).
//
// Same class shape as src/decoder_template/CGPUDecoder.h:19-35:
//     CGPUDecoder(size_t nb_frames /* words */, size_t n, size_t k, size_t m);
//     virtual void initialize() = 0;
//     virtual void decode(float var_nodes[], int Rprime_fix[], int iterations) = 0;
// plus the five subclasses (src/decoder_{ms,oms,nms,2nms,oms_v2}/CGPU_Decoder_*.h),
// CTrame's pinned frame buffers (src/trame/CTrame.h) and CTimer
// (src/timer/CTimer.h).  The reference's code table is compile-time; here
// the constructor also takes the code (ldpc_code*).  Errors throw
// ldpc::Error instead of printing and calling exit(0)
// (src/custom_api/custom_cuda.cu:5-17).
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>

#include "ldpc_mi355x.h"

namespace ldpc {

struct Error : std::runtime_error {
    explicit Error(const std::string &what) : std::runtime_error(what + ": " + ldpc_last_error()) {}
};

inline void check(int rc, const char *what) {
    if (rc != LDPC_OK) throw Error(what);
}

}  // namespace ldpc

class CGPUDecoder {
protected:
    ldpc_decoder *dec_ = nullptr;
    size_t nb_frames;  // 32-bit words, 4 frames each
    size_t sz_nodes, sz_checks, sz_msgs;

    CGPUDecoder(size_t _nb_frames, size_t n, size_t k, size_t m, const ldpc_code *code, int variant,
                const ldpc_tuning *tuning)
        : nb_frames(_nb_frames), sz_nodes(n * _nb_frames), sz_checks(k * _nb_frames), sz_msgs(m * _nb_frames) {
        int cn = 0, ck = 0, cm = 0;
        ldpc::check(ldpc_code_info(code, &cn, &ck, &cm, nullptr, nullptr, nullptr), "ldpc_code_info");
        if ((size_t)cn != n || (size_t)ck != k || (size_t)cm != m)
            throw std::invalid_argument("CGPUDecoder: (n, k, m) does not match the code");
        dec_ = ldpc_decoder_create(code, _nb_frames, variant, tuning);
        if (!dec_) throw ldpc::Error("ldpc_decoder_create");
    }

public:
    CGPUDecoder(const CGPUDecoder &) = delete;
    CGPUDecoder &operator=(const CGPUDecoder &) = delete;
    virtual ~CGPUDecoder() { ldpc_decoder_destroy(dec_); }

    virtual void initialize() { ldpc::check(ldpc_decoder_initialize(dec_), "initialize"); }

    // var_nodes: 4W*N packed int8 LLRs (frame-major; device layout for
    // OMS_v2); Rprime_fix: 4W*N bytes of hard decisions.
    virtual void decode(float var_nodes[], int Rprime_fix[], int nombre_iterations) {
        ldpc::check(ldpc_decoder_decode(dec_, reinterpret_cast<const int8_t *>(var_nodes),
                                        reinterpret_cast<uint8_t *>(Rprime_fix), nombre_iterations),
                    "decode");
    }

    // device-buffer form, asynchronous on `stream` (a hipStream_t, or null)
    void decode_device(const int8_t *d_in, uint8_t *d_out, int iterations, void *stream = nullptr) {
        ldpc::check(ldpc_decoder_decode_device(dec_, d_in, d_out, iterations, stream), "decode_device");
    }

    ldpc_decoder *handle() const { return dec_; }
};

#define LDPC_DECODER_CLASS(NAME, VARIANT)                                                              \
    class NAME : public CGPUDecoder {                                                                  \
    public:                                                                                            \
        NAME(size_t nb_frames, size_t n, size_t k, size_t m, const ldpc_code *code,                    \
             const ldpc_tuning *tuning = nullptr)                                                      \
            : CGPUDecoder(nb_frames, n, k, m, code, VARIANT, tuning) {}                                \
    };

LDPC_DECODER_CLASS(CGPU_Decoder_MS_SIMD, LDPC_MS)      // src/decoder_ms/CGPU_Decoder_MS_SIMD.h
LDPC_DECODER_CLASS(CGPU_Decoder_OMS_SIMD, LDPC_OMS)    // src/decoder_oms/CGPU_Decoder_OMS_SIMD.h
LDPC_DECODER_CLASS(CGPU_Decoder_NMS_SIMD, LDPC_NMS)    // src/decoder_nms/CGPU_Decoder_NMS_SIMD.h
LDPC_DECODER_CLASS(CGPU_Decoder_2NMS_SIMD, LDPC_2NMS)  // src/decoder_2nms/CGPU_Decoder_2NMS_SIMD.h
LDPC_DECODER_CLASS(CGPU_Decoder_MS_SIMD_v2, LDPC_OMS_V2)  // src/decoder_oms_v2/CGPU_Decoder_MS_SIMD_v2.h
#undef LDPC_DECODER_CLASS

// Pinned host frame buffers (src/trame/CTrame.cpp:26-45): width = N,
// height = #checks, frame = words.  nb_vars() = N - #checks as in the reference.
class CTrame {
    int _width, _height, _frame;
    float *t_noise_data = nullptr;
    int *t_decode_data = nullptr;

public:
    CTrame(int width, int height, int frame) : _width(width), _height(height), _frame(frame) {
        const size_t words = (size_t)width * (size_t)frame + 4;
        t_noise_data = static_cast<float *>(ldpc_host_alloc_pinned(words * sizeof(float)));
        t_decode_data = static_cast<int *>(ldpc_host_alloc_pinned(words * sizeof(int)));
        if (!t_noise_data || !t_decode_data) {
            ldpc_host_free_pinned(t_noise_data);
            ldpc_host_free_pinned(t_decode_data);
            throw ldpc::Error("CTrame pinned allocation");
        }
    }
    CTrame(const CTrame &) = delete;
    CTrame &operator=(const CTrame &) = delete;
    ~CTrame() {
        ldpc_host_free_pinned(t_noise_data);
        ldpc_host_free_pinned(t_decode_data);
    }
    int nb_vars() { return _width - _height; }
    int nb_frames() { return _frame; }
    int nb_checks() { return _height; }
    int nb_data() { return _width; }
    float *get_t_noise_data() { return t_noise_data; }
    int *get_t_decode_data() { return t_decode_data; }
};

// hipEvent stopwatch (src/timer/CTimer.h:18-48)
class CTimer {
    ldpc_timer *t_;
    bool running_ = false;

public:
    explicit CTimer(bool start = false, void *stream = nullptr) : t_(ldpc_timer_create(stream)) {
        if (!t_) throw ldpc::Error("ldpc_timer_create");
        if (start) this->start();
    }
    CTimer(const CTimer &) = delete;
    CTimer &operator=(const CTimer &) = delete;
    ~CTimer() { ldpc_timer_destroy(t_); }
    void start() {
        ldpc::check(ldpc_timer_start(t_), "timer start");
        running_ = true;
    }
    void stop() {
        ldpc::check(ldpc_timer_stop(t_), "timer stop");
        running_ = false;
    }
    double get_time_ms() {
        if (running_) stop();
        return ldpc_timer_elapsed_ms(t_);
    }
    double get_time_sec() { return get_time_ms() / 1000.0; }
};

// ---------------------------------------------------------------------------
// Batched decoding with per-frame stop and slot refill (no counterpart in
// the reference): RAII over ldpc_pool_* / ldpc_mixed_*.
// ---------------------------------------------------------------------------
class CFrameQueueDecoder {
    ldpc_pool *p_;

public:
    CFrameQueueDecoder(const ldpc_code *code, int variant, int max_iterations, int stop_rule = LDPC_STOP_SYNDROME,
                       const ldpc_tuning *tuning = nullptr)
        : p_(ldpc_pool_create(code, variant, max_iterations, stop_rule, tuning)) {
        if (!p_) throw ldpc::Error("ldpc_pool_create");
    }
    CFrameQueueDecoder(const CFrameQueueDecoder &) = delete;
    CFrameQueueDecoder &operator=(const CFrameQueueDecoder &) = delete;
    ~CFrameQueueDecoder() { ldpc_pool_destroy(p_); }

    // the fastest tuning for a code (timed on the GPU)
    static ldpc_tuning autotune(const ldpc_code *code, int variant, int max_iterations, double ebn0_db,
                                int quant = 0, uint64_t sample_frames = 32768, int stop_rule = LDPC_STOP_SYNDROME) {
        ldpc_tuning t;
        ldpc::check(ldpc_pool_autotune(code, variant, max_iterations, stop_rule, ebn0_db, quant, sample_frames, &t,
                                       nullptr, nullptr),
                    "ldpc_pool_autotune");
        return t;
    }
    // int8 table [n_iters][n_degs] for LDPC_LOMS / LDPC_LNMS
    void set_table(const int8_t *table, int n_iters, int n_degs) {
        ldpc::check(ldpc_pool_set_table(p_, table, n_iters, n_degs), "ldpc_pool_set_table");
    }
    // LDPC_ORDER_INPUT or LDPC_ORDER_LEAST_RELIABLE_FIRST (same results, other queue order)
    void set_order(int order) { ldpc::check(ldpc_pool_set_order(p_, order), "ldpc_pool_set_order"); }
    // frame-major device buffers; d_iters may be null
    void decode_device(const int8_t *d_in, uint64_t frames, uint8_t *d_out, int *d_iters = nullptr,
                       void *stream = nullptr, int workgroups = 0) {
        ldpc::check(ldpc_pool_decode(p_, d_in, frames, d_out, d_iters, workgroups, stream), "ldpc_pool_decode");
    }
    // the same step captured in a hipGraph once per buffer set and replayed
    void decode_graph(const int8_t *d_in, uint64_t frames, uint8_t *d_out, int *d_iters = nullptr,
                      void *stream = nullptr, int workgroups = 0) {
        ldpc::check(ldpc_pool_decode_graph(p_, d_in, frames, d_out, d_iters, workgroups, stream),
                    "ldpc_pool_decode_graph");
    }
    ldpc_sim_stats simulate(double ebn0_db, uint64_t frames, uint64_t seed = 1, uint32_t snr_idx = 0, int quant = 0,
                            uint64_t batch_frames = 0) {
        ldpc_sim_stats s;
        ldpc::check(ldpc_pool_simulate(p_, ebn0_db, quant, seed, snr_idx, frames, batch_frames, &s),
                    "ldpc_pool_simulate");
        return s;
    }
    void *stream() const { return ldpc_pool_stream(p_); }
    ldpc_pool *handle() const { return p_; }
};

class CMixedDecoder {
    ldpc_mixed *m_;

public:
    CMixedDecoder(const ldpc_code *const *codes, int n_codes, int variant, int max_iterations,
                  const ldpc_tuning *tunings = nullptr, uint64_t max_frames = 0,
                  int stop_rule = LDPC_STOP_SYNDROME)
        : m_(ldpc_mixed_create(codes, n_codes, variant, max_iterations, stop_rule, tunings, max_frames)) {
        if (!m_) throw ldpc::Error("ldpc_mixed_create");
    }
    CMixedDecoder(const CMixedDecoder &) = delete;
    CMixedDecoder &operator=(const CMixedDecoder &) = delete;
    ~CMixedDecoder() { ldpc_mixed_destroy(m_); }

    void set_table(int code_index, const int8_t *table, int n_iters, int n_degs) {
        ldpc::check(ldpc_mixed_set_table(m_, code_index, table, n_iters, n_degs), "ldpc_mixed_set_table");
    }
    // frame f: code d_code_ids[f], LLRs at d_in + d_offsets[f]; decisions at d_out + d_offsets[f]
    void decode(const int8_t *d_in, const uint64_t *d_offsets, const uint8_t *d_code_ids, uint64_t frames,
                uint8_t *d_out, int *d_iters, bool use_graph = true, void *stream = nullptr) {
        ldpc::check(ldpc_mixed_decode(m_, d_in, d_offsets, d_code_ids, frames, d_out, d_iters, use_graph ? 1 : 0,
                                      stream),
                    "ldpc_mixed_decode");
    }
    ldpc_mixed *handle() const { return m_; }
};
