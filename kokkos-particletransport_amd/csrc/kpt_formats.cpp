// kpt_formats.cpp — host-only CSV outputs of the driver, byte-compatible with
// write_momentum_histogram (src/main.cpp:1944-1963) and write_summary_header /
// write_summary_row (src/main.cpp:2207-2235). Those stream doubles with
// std::setprecision(16) in the default float field, which is printf's %.16g.
#include <cmath>
#include <cstdio>
#include <string>

#include "../../include/kpt.h"

namespace {

void append_g16(std::string& s, double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.16g", v);
    s += b;
}

void append_u64(std::string& s, unsigned long long v) {
    char b[32];
    std::snprintf(b, sizeof(b), "%llu", v);
    s += b;
}

kpt_status copy_out(const std::string& s, char* buf, size_t len) {
    if (!buf || len < s.size() + 1) return KPT_ERR_ARG;
    std::snprintf(buf, len, "%s", s.c_str());
    return KPT_OK;
}

}  // namespace

extern "C" {

kpt_status kpt_write_histogram_csv(const char* path, int32_t bins, double pmin, double pmax,
                                   const double* weights) {
    if (!path || !weights || bins <= 0 || !(pmin > 0.0) || !(pmax > pmin)) return KPT_ERR_ARG;
    std::string s = "bin,p_left,p_right,weight\n";
    const double dlogp = (std::log(pmax) - std::log(pmin)) / static_cast<double>(bins);
    // The reference is built with -O3 -march=native (README.en.md:76-81), where
    // GCC contracts ln(pmin) + b * dlogp into one fused multiply-add; the edges of
    // its tracked CSV carry that rounding, so the contraction is written out here.
    for (int b = 0; b < bins; ++b) {
        const double p_left = std::exp(std::fma(static_cast<double>(b), dlogp, std::log(pmin)));
        const double p_right = std::exp(std::fma(static_cast<double>(b + 1), dlogp, std::log(pmin)));
        append_u64(s, static_cast<unsigned long long>(b));
        s += ',';
        append_g16(s, p_left);
        s += ',';
        append_g16(s, p_right);
        s += ',';
        append_g16(s, weights[b]);
        s += '\n';
    }
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return KPT_ERR_IO;
    const size_t n = std::fwrite(s.data(), 1, s.size(), f);
    const int rc = std::fclose(f);
    return (n == s.size() && rc == 0) ? KPT_OK : KPT_ERR_IO;
}

kpt_status kpt_format_summary_header(char* buf, size_t len) {
    return copy_out(
        "frame,time,stored,active,inactive,weight_sum,p_min,p_max,p_avg,"
        "dt_min,dt_max,dt_avg,max_split_level,injected,split,"
        "skipped_split_capacity,read_seconds,inject_seconds,move_seconds,"
        "split_seconds,total_frame_seconds,total_elapsed_seconds\n",
        buf, len);
}

kpt_status kpt_format_summary_row(char* buf, size_t len, int32_t frame, double time, const kpt_summary* sm,
                                  const kpt_frame_timing* tm, double total_frame_seconds,
                                  double total_elapsed_seconds) {
    if (!sm || !tm) return KPT_ERR_ARG;
    std::string s;
    char b[32];
    std::snprintf(b, sizeof(b), "%d", frame);
    s += b;
    s += ',';
    append_g16(s, time);
    s += ',';
    append_u64(s, sm->stored);
    s += ',';
    append_u64(s, sm->active);
    s += ',';
    append_u64(s, sm->inactive);
    s += ',';
    for (double v : {sm->weight_sum, sm->p_min, sm->p_max, sm->p_avg, sm->dt_min, sm->dt_max, sm->dt_avg}) {
        append_g16(s, v);
        s += ',';
    }
    std::snprintf(b, sizeof(b), "%d", sm->max_split_level);
    s += b;
    s += ',';
    append_u64(s, tm->injected);
    s += ',';
    append_u64(s, tm->split);
    s += ',';
    append_u64(s, tm->skipped_split_capacity);
    s += ',';
    const double tail[6] = {tm->read_seconds, tm->inject_seconds, tm->move_seconds, tm->split_seconds,
                            total_frame_seconds, total_elapsed_seconds};
    for (int k = 0; k < 6; ++k) {
        append_g16(s, tail[k]);
        s += k < 5 ? ',' : '\n';
    }
    return copy_out(s, buf, len);
}

}  // extern "C"
