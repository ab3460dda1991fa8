// ldpc_sim -- the Eb/N0 sweep of src/main.cpp on the C ABI of ldpc_mi355x.h.
//
// Same options and loop as the reference driver (src/main.cpp:121-335):
//   -min/-max/-pas   sweep range and step (dB)        -iter  iterations
//   -fer             frame errors per point           -timer stop after this many seconds (-1: none)
//   -qef BER         stop the sweep once a point's BER falls below BER (quasi error free)
//   -n W             words per decode (4W frames)     -Eb/N0 | -Es/N0   SNR convention
//   -bpsk | -qpsk    accepted; Gray-mapped QPSK gives every bit the BPSK channel at the same Eb/N0
//   -fMS -MS -OMS -NMS -2NMS -xMS                     the decoder (fMS is MS, xMS is OMS_v2)
//   -info            device and library information
// and, beyond the reference:
//   -code 3_6 | PATH  the (3,6) N=4000 code (default) or a table file in the
//                     reference's matrix format (-N gives its length)
//   -quant 0|1        LLR quantizer: 0 = sat127(round(16y/sigma^2)), 1 = the
//                     reference's trunc(8y) clamped to +-31 (src/awgn_channel/CChanel_AWGN_SIMD.cu:17-28)
//   -pool             the frame-queue decoder (per-frame syndrome stop, -iter as the limit)
//   -order            with -pool: queue each batch's frames least reliable first
//   -maxframes F      frames per point at most (0: no cap)
//   -seed S
// Per point: the device channel draws 4W frames per batch (a new Philox
// stream per batch), the decoder runs, the device counters count bit and
// frame errors in the first N - #checks positions; batches repeat until
// the frame errors reach -fer (CErrorAnalyzer::fe_limit_achieved,
// src/ber_analyzer/CErrorAnalyzer.cpp:113-116), the frame cap or the timer.
// Printed per point: the reference's final-report fields plus information
// throughput from the decode time alone and the algorithmic HBM rate.
#include "ldpc_mi355x.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

struct Options {
    double min_db = 0.50, max_db = 3.01, step_db = 0.10;
    int iters = 20, fe_limit = 200, timer_s = -1;
    bool qef = false;
    double qef_ber = 1e-7;
    bool es_n0 = false;
    size_t words = 1024;
    int variant = LDPC_MS;
    const char *name = "fMS";
    std::string code = "3_6";
    int n_vars = 0;
    int quant = 0;
    bool pool = false;
    bool order = false;
    unsigned long long max_frames = 0;
    unsigned long long seed = 1;
};

void usage() {
    std::printf("usage: ldpc_sim [-min dB] [-max dB] [-pas dB] [-iter n] [-fer n] [-timer s] [-qef ber]\n"
                "                [-n words] [-Eb/N0|-Es/N0] [-bpsk|-qpsk] [-fMS|-MS|-OMS|-NMS|-2NMS|-xMS]\n"
                "                [-code 3_6|PATH] [-N n] [-quant 0|1] [-pool [-order]] [-maxframes F] [-seed S] [-info]\n");
}

int fail(const char *what) {
    std::fprintf(stderr, "(EE) %s: %s\n", what, ldpc_last_error());
    return 1;
}

void show_time(unsigned long s) { std::printf("%2.2luh%2.2lu'%2.2lu", s / 3600, (s / 60) % 60, s % 60); }

}  // namespace

int main(int argc, char **argv) {
    Options o;
    for (int p = 1; p < argc; p++) {
        const char *a = argv[p];
        const bool has = p + 1 < argc;
        if (!std::strcmp(a, "-min") && has) o.min_db = std::atof(argv[++p]);
        else if (!std::strcmp(a, "-max") && has) o.max_db = std::atof(argv[++p]);
        else if (!std::strcmp(a, "-pas") && has) o.step_db = std::atof(argv[++p]);
        else if (!std::strcmp(a, "-timer") && has) o.timer_s = std::atoi(argv[++p]);
        else if (!std::strcmp(a, "-iter") && has) o.iters = std::atoi(argv[++p]);
        else if (!std::strcmp(a, "-fer") && has) o.fe_limit = std::atoi(argv[++p]);
        else if (!std::strcmp(a, "-qef") && has) { o.qef = true; o.qef_ber = std::atof(argv[++p]); }
        else if (!std::strcmp(a, "-bpsk") || !std::strcmp(a, "-qpsk")) {}
        else if (!std::strcmp(a, "-Eb/N0")) o.es_n0 = false;
        else if (!std::strcmp(a, "-Es/N0")) o.es_n0 = true;
        else if (!std::strcmp(a, "-n") && has) o.words = (size_t)std::atoll(argv[++p]);
        else if (!std::strcmp(a, "-fMS")) { o.variant = LDPC_MS; o.name = "fMS"; }
        else if (!std::strcmp(a, "-MS")) { o.variant = LDPC_MS; o.name = "MS"; }
        else if (!std::strcmp(a, "-OMS")) { o.variant = LDPC_OMS; o.name = "OMS"; }
        else if (!std::strcmp(a, "-NMS")) { o.variant = LDPC_NMS; o.name = "NMS"; }
        else if (!std::strcmp(a, "-2NMS")) { o.variant = LDPC_2NMS; o.name = "2NMS"; }
        else if (!std::strcmp(a, "-xMS")) { o.variant = LDPC_OMS_V2; o.name = "xMS"; }
        else if (!std::strcmp(a, "-code") && has) o.code = argv[++p];
        else if (!std::strcmp(a, "-N") && has) o.n_vars = std::atoi(argv[++p]);
        else if (!std::strcmp(a, "-quant") && has) o.quant = std::atoi(argv[++p]);
        else if (!std::strcmp(a, "-pool")) o.pool = true;
        else if (!std::strcmp(a, "-order")) o.order = true;
        else if (!std::strcmp(a, "-maxframes") && has) o.max_frames = std::strtoull(argv[++p], nullptr, 10);
        else if (!std::strcmp(a, "-seed") && has) o.seed = std::strtoull(argv[++p], nullptr, 10);
        else if (!std::strcmp(a, "-info")) {
            std::printf("(II) %s, %d device(s)\n", ldpc_version(), ldpc_device_count());
            return 0;
        } else {
            std::printf("(EE) Unknown argument (%d) => [%s]\n", p, a);
            usage();
            return 2;
        }
    }
    if (o.iters < 1 || o.fe_limit < 1 || o.words < 1 || o.step_db <= 0.0 || (o.quant != 0 && o.quant != 1)) {
        usage();
        return 2;
    }
    if (ldpc_device_count() < 1) {
        std::fprintf(stderr, "(EE) no gfx950 device\n");
        return 1;
    }

    ldpc_code *code = nullptr;
    if (o.code == "3_6") {
        ldpc_code *raw = ldpc_code_regular(4000, 3, 6, 1);
        if (raw) code = ldpc_code_reorder_by_colour(raw);
        ldpc_code_destroy(raw);
    } else {
        code = ldpc_code_load_table(o.code.c_str(), o.n_vars);
    }
    if (!code) return fail("code");
    int N = 0, C = 0, E = 0, max_deg = 0, levels = 0, q1 = 0;
    if (ldpc_code_info(code, &N, &C, &E, &max_deg, &levels, &q1) != LDPC_OK) return fail("code info");
    const int K = N - C;
    const double rate = (double)K / N;

    ldpc_decoder *dec = nullptr;
    ldpc_pool *pool = nullptr;
    if (o.pool) {
        pool = ldpc_pool_create(code, o.variant, o.iters, LDPC_STOP_SYNDROME, nullptr);
        if (!pool) return fail("pool");
        if (o.order && ldpc_pool_set_order(pool, LDPC_ORDER_LEAST_RELIABLE_FIRST) != LDPC_OK) return fail("order");
    } else {
        dec = ldpc_decoder_create(code, o.words, o.variant, nullptr);
        if (!dec || ldpc_decoder_initialize(dec) != LDPC_OK) return fail("decoder");
    }
    const unsigned long long batch = 4ull * o.words;

    std::printf("(II) Code LDPC (N, K)     : (%d,%d), %d edges, max check degree %d, %d levels\n", N, K, E, max_deg,
                levels);
    std::printf("(II) Rendement du code    : %.3f\n", rate);
    std::printf("(II) # ITERATIONs du CODE : %d%s\n", o.iters, o.pool ? " (limit; per-frame syndrome stop)" : "");
    std::printf("(II) FER LIMIT FOR SIMU   : %d\n", o.fe_limit);
    std::printf("(II) SIMULATION  RANGE    : [%.2f, %.2f], STEP = %.2f\n", o.min_db, o.max_db, o.step_db);
    std::printf("(II) MODE EVALUATION      : %s\n", o.es_n0 ? "Es/N0" : "Eb/N0");
    std::printf("(II) MIN-SUM ALGORITHM    : %s%s\n", o.name, o.pool ? " (frame queue)" : "");
    std::printf("(II) FRAMES PER BATCH     : %llu\n", batch);
    std::fflush(stdout);

    const auto t_start = std::chrono::steady_clock::now();
    auto elapsed_s = [&] {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    };
    int rc = 0;
    uint32_t point = 0;
    for (double snr = o.min_db; snr <= o.max_db; snr += o.step_db, point++) {
        // Es/N0 = Eb/N0 + 10 log10(R) for one coded bit per symbol dimension
        const double ebn0 = o.es_n0 ? snr - 10.0 * std::log10(rate) : snr;
        const auto t_point = std::chrono::steady_clock::now();
        unsigned long long frames = 0, fe = 0, be = 0, k = 0;
        double decode_ms = 0.0, it_sum = 0.0;
        bool timed_out = false;
        while (fe < (unsigned long long)o.fe_limit && (!o.max_frames || frames < o.max_frames)) {
            ldpc_sim_stats st;
            const unsigned long long seed = o.seed + 0x9E3779B97F4A7C15ull * (k + 1);
            const int r = o.pool ? ldpc_pool_simulate(pool, ebn0, o.quant, seed, point, batch, 0, &st)
                                 : ldpc_simulate(dec, ebn0, o.quant, seed, point, batch, o.iters, &st);
            if (r != LDPC_OK) {
                rc = fail("simulate");
                break;
            }
            frames += st.frames;
            fe += st.frame_errors;
            be += st.bit_errors;
            decode_ms += st.decode_ms;
            it_sum += st.avg_iterations * (double)st.frames;
            k++;
            if (o.timer_s != -1 && elapsed_s() >= o.timer_s) {
                timed_out = true;
                break;
            }
        }
        if (rc) break;
        const double ber = (double)be / (double)frames / (double)K;
        const double fer = (double)fe / (double)frames;
        const double avg_it = it_sum / (double)frames;
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_point).count();
        const double info_gbps = (double)frames * K / (decode_ms * 1e-3) / 1e9;
        // algorithmic bytes: per frame-iteration 2E + 2N (messages and posterior
        // read and written), plus the LLRs in and the decisions out
        const double alg = (double)frames * (avg_it * (2.0 * E + 2.0 * N) + 2.0 * N);
        std::printf("SNR = %.2f | BER =  %2.3e | FER =  %2.3e | BPS =  %2.2f | MATRICES = %10llu| FE = %llu | BE = %llu | "
                    "BE/FE = %.1f | AVG IT = %.2f | INFO = %.3f Gbit/s | HBM (alg.) = %.0f GB/s | RUNTIME = ",
                    snr, ber, fer, (double)frames * K / (secs > 0 ? secs : 1.0) / 1e6, frames, fe, be,
                    fe ? (double)be / (double)fe : 0.0, avg_it, info_gbps, alg / (decode_ms * 1e-3) / 1e9);
        show_time((unsigned long)secs + 1);
        std::printf("\n");
        std::fflush(stdout);
        if (timed_out) {
            std::printf("(II) THE SIMULATION HAS STOP DUE TO THE (USER) TIME CONTRAINT.\n");
            break;
        }
        if (o.qef && ber < o.qef_ber) {
            std::printf("(II) THE SIMULATION HAS STOP DUE TO THE (USER) QUASI-ERROR FREE CONTRAINT.\n");
            break;
        }
    }
    if (pool) ldpc_pool_destroy(pool);
    if (dec) ldpc_decoder_destroy(dec);
    ldpc_code_destroy(code);
    return rc;
}
