// kpt_run.cpp — the reconnection calibration driver on the C ABI of include/kpt.h.
//
// Mirrors Reconnection2DCalibration::run / run_transport_loop
// (src/main.cpp:2253-2429, :2615-2690): the same command-line controls and
// profiles (:2444-2611, :528-590), the same validation messages (:183-241),
// the same outputs in --output-dir (summary.csv, run.log,
// momentum_histogram_XXXXX.csv, particles_XXXXX.bin) and the same per-frame
// order: [load upper frame] -> inject -> move -> split -> [summary row] ->
// [histogram] -> [snapshot] -> [walltime check]. The push runs on the GPU
// through kpt_push_interval; --precision, --push-chunks, the closure options,
// --resumable-snapshots, --allow-stretched-fields, --synchronous-field-reads and
// --device select this program's extensions.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kpt.h"

namespace {

enum class Transport { Parker, Focused };

// Reconnection2DSettings (src/main.cpp:114-154) plus the extensions.
struct RunSettings {
    std::string profile_name{"particle-64000"};
    std::string field_dir{"cmake-build-debug/field_fortran_check"};
    std::string output_dir{"cmake-build-debug/reconnection_particle_64000"};
    std::string restart_particle_snapshot_path{};
    Transport transport{Transport::Parker};
    int start_frame{0};
    int end_frame{2};
    uint64_t particles_per_frame{64000};
    uint64_t rank_scale{1};
    uint64_t particle_capacity{1000000};
    uint64_t seed{114514};
    int diagnostic_interval{1};
    int histogram_interval{1};
    int particle_snapshot_interval{1};
    int max_particle_steps_per_interval{200000};
    bool split_particles{true};
    bool time_interpolation{true};
    bool overwrite_output{false};
    double walltime_limit_seconds{0.0};
    double walltime_reserve_seconds{0.0};
    kpt_host_settings host{};
    // extensions
    int precision{KPT_PREC_FP64};
    int push_chunks{8};
    int closure{KPT_CLOSURE_REDUCED};
    double qlt_sigma2{0.1};
    double qlt_correlation_length{0.02};
    double qlt_spectral_index{5.0 / 3.0};
    uint32_t mlp_train_steps{8000};
    int32_t mlp_activation{KPT_MLP_SILU};
    int device{0};
    bool resumable_snapshots{false};
    bool allow_stretched_fields{false};
    bool stream_fields{true};

    int interval_count() const { return end_frame - start_frame; }
    bool walltime_limit_enabled() const { return walltime_limit_seconds > 0.0; }
    bool restart_enabled() const { return !restart_particle_snapshot_path.empty(); }
    double walltime_stop_seconds() const { return std::max(0.0, walltime_limit_seconds - walltime_reserve_seconds); }

    // Reconnection2DSettings::validate (src/main.cpp:183-241)
    void validate() const {
        const kpt_host_settings& h = host;
        if (end_frame <= start_frame)
            throw std::runtime_error("Reconnection2DSettings: end frame must exceed start frame.");
        if (start_frame < 0) throw std::runtime_error("Reconnection2DSettings: start frame must be non-negative.");
        if (restart_enabled() && !std::filesystem::exists(restart_particle_snapshot_path))
            throw std::runtime_error("Reconnection2DSettings: restart particle snapshot does not exist.");
        if (restart_enabled() && overwrite_output)
            throw std::runtime_error("Reconnection2DSettings: --overwrite-output cannot be combined with restart.");
        if (particles_per_frame == 0 || rank_scale == 0 || particle_capacity == 0)
            throw std::runtime_error("Reconnection2DSettings: particle counts and capacity must be positive.");
        if (diagnostic_interval <= 0 || particle_snapshot_interval < 0 || max_particle_steps_per_interval <= 0)
            throw std::runtime_error(
                "Reconnection2DSettings: diagnostic interval and step guard must be positive, and particle "
                "snapshot interval must be non-negative.");
        if (!std::isfinite(walltime_limit_seconds) || walltime_limit_seconds < 0.0 ||
            !std::isfinite(walltime_reserve_seconds) || walltime_reserve_seconds < 0.0 ||
            (walltime_limit_seconds <= 0.0 && walltime_reserve_seconds > 0.0))
            throw std::runtime_error(
                "Reconnection2DSettings: walltime limit must be non-negative, and reserve requires a positive "
                "walltime limit.");
        if (!std::isfinite(h.mhd_dt) || h.mhd_dt <= 0.0 || !std::isfinite(h.p0) || h.p0 <= 0.0 ||
            !std::isfinite(h.pmin) || h.pmin <= 0.0 || !std::isfinite(h.pmax) || h.pmax <= h.pmin ||
            !std::isfinite(h.gamma_turb) || !std::isfinite(h.kpara0) || h.kpara0 <= 0.0 ||
            !std::isfinite(h.kperp_over_kpara) || h.kperp_over_kpara < 0.0 || !std::isfinite(h.dt_min_rel) ||
            h.dt_min_rel <= 0.0 || !std::isfinite(h.dt_max_rel) || h.dt_max_rel < h.dt_min_rel ||
            !std::isfinite(h.drift_param1) || !std::isfinite(h.drift_param2) || !std::isfinite(h.charge) ||
            h.charge == 0.0 || !std::isfinite(h.particle_v0) || h.particle_v0 <= 0.0 || !std::isfinite(h.duu0) ||
            h.duu0 < 0.0 || !std::isfinite(h.split_ratio) || h.split_ratio <= 1.0 ||
            !std::isfinite(h.pmin_split_over_p0) || h.pmin_split_over_p0 <= 0.0 || !std::isfinite(h.mu_max) ||
            h.mu_max <= 0.0)
            throw std::runtime_error("Reconnection2DSettings: invalid physical parameters.");
        if (push_chunks < 0 || push_chunks > 4096)
            throw std::runtime_error("kpt_run: --push-chunks must be in [0, 4096].");
        if (closure == KPT_CLOSURE_MLP && precision != KPT_PREC_FP32)
            throw std::runtime_error("kpt_run: the learned pitch-angle closure runs with --precision fp32.");
        if (closure != KPT_CLOSURE_REDUCED && transport != Transport::Focused)
            throw std::runtime_error("kpt_run: --closure qlt / mlp needs --transport focused.");
    }
};

// apply_run_profile (src/main.cpp:528-590)
void apply_profile(RunSettings& s, const std::string& name) {
    const kpt_host_settings host = s.host;
    s = RunSettings{};
    s.host = host;
    s.profile_name = name;
    if (name == "particle-64000" || name == "particle-test") {
        s.output_dir = "cmake-build-debug/reconnection_particle_64000";
        s.end_frame = 2;
        s.particles_per_frame = 64000;
        s.particle_capacity = 1000000;
        return;
    }
    if (name == "smoke") {
        s.output_dir = "cmake-build-debug/reconnection_smoke";
        s.end_frame = 1;
        s.particles_per_frame = 1024;
        s.particle_capacity = 8192;
        return;
    }
    if (name == "fortran-rank") {
        s.field_dir = "cmake-build-debug/field";
        s.output_dir = "cmake-build-debug/reconnection_fortran_rank";
        s.end_frame = 200;
        s.particles_per_frame = 1600;
        s.particle_capacity = 1000000;
        s.histogram_interval = 10;
        s.particle_snapshot_interval = 10;
        return;
    }
    if (name == "fortran-global" || name == "full") {
        s.field_dir = "cmake-build-debug/field";
        s.output_dir = "cmake-build-debug/reconnection_fortran_global";
        s.end_frame = 200;
        s.particles_per_frame = 1600;
        s.rank_scale = 16;
        s.particle_capacity = 16000000;
        s.histogram_interval = 10;
        s.particle_snapshot_interval = 10;
        return;
    }
    throw std::runtime_error("Unknown run profile: " + name + ".");
}

std::string next_value(int argc, char** argv, int& i, const std::string& opt) {
    if (i + 1 >= argc) throw std::runtime_error("Missing value for " + opt + ".");
    return argv[++i];
}

std::string frame_path(const std::string& dir, const char* prefix, int frame, const char* suffix) {
    std::ostringstream p;
    p << dir << "/" << prefix << std::setw(5) << std::setfill('0') << frame << suffix;
    return p.str();
}

// infer_particle_snapshot_frame: particles_XXXXX.bin -> XXXXX, else -1.
int snapshot_frame(const std::string& path) {
    const std::string name = std::filesystem::path(path).filename().string();
    const std::string pre = "particles_", suf = ".bin";
    if (name.size() <= pre.size() + suf.size() || name.compare(0, pre.size(), pre) != 0 ||
        name.compare(name.size() - suf.size(), suf.size(), suf) != 0)
        return -1;
    const std::string digits = name.substr(pre.size(), name.size() - pre.size() - suf.size());
    for (char c : digits)
        if (c < '0' || c > '9') return -1;
    return std::stoi(digits);
}

// --help: the option names are the reference driver's command line
// (src/main.cpp:2444-2611); the grouping and descriptions are this program's.
struct HelpEntry {
    const char* option;
    const char* text;
};

void print_help() {
    static const struct {
        const char* title;
        std::vector<HelpEntry> entries;
    } groups[] = {
        {"Run and transport",
         {{"--profile NAME", "preset: particle-64000 (default), smoke, fortran-rank, fortran-global, full"},
          {"--transport parker|focused", "transport equation (default parker)"},
          {"--parker-transport", "same as --transport parker"},
          {"--focused-transport", "same as --transport focused"},
          {"--device N", "GPU ordinal (default 0)"}}},
        {"Frames and files",
         {{"--field-dir PATH", "compact-field frames (fieldXXXXX.bin) to read"},
          {"--output-dir PATH", "where CSVs, run.log and snapshots go"},
          {"--start-frame N", "first frame (default 0)"},
          {"--end-frame N", "last frame read"},
          {"--frames N", "end frame = start frame + N"},
          {"--dt-out X", "time between frames (default 0.1)"},
          {"--overwrite-output", "a fresh run may replace existing outputs"},
          {"--restart-particle-snapshot PATH", "continue from particles_XXXXX.bin at frame XXXXX"},
          {"--resumable-snapshots", "snapshots keep ids and split baselines (bitwise restarts)"},
          {"--allow-stretched-fields", "accept non-uniform faces (local widths in dt)"},
          {"--synchronous-field-reads", "read each frame between intervals, not during the push"}}},
        {"Particles",
         {{"--particles-per-frame N", "injected per frame (profile value otherwise)"},
          {"--rank-scale N", "multiplies the per-frame injection"},
          {"--capacity N", "particle storage (profile value otherwise)"},
          {"--seed N", "random seed (default 114514)"},
          {"--particle-v0 X", "focused transport: speed at p0"},
          {"--duu0 X", "focused transport: D_mumu scale"},
          {"--no-split", "no momentum splitting"},
          {"--split-ratio X", "momentum ratio between split levels"},
          {"--pmin-split-over-p0 X", "first split momentum, in units of p0"}}},
        {"Push",
         {{"--max-steps-per-interval N", "sub-step cap per frame interval (default 200000)"},
          {"--no-time-interp", "sample the lower frame only"},
          {"--precision fp64|fp32", "arithmetic of the push (default fp64)"},
          {"--push-chunks N", "cell sorts per interval (default 8)"},
          {"--closure reduced|qlt|mlp", "focused pitch-angle closure (default reduced)"},
          {"--qlt-sigma2 X", "slab turbulence level (default 0.1)"},
          {"--qlt-correlation-length X", "slab correlation length (default 0.02)"},
          {"--qlt-spectral-index X", "slab spectral index (default 5/3)"},
          {"--mlp-train-steps N", "training steps of the learned closure (default 8000)"},
          {"--mlp-activation silu|hard-swish", "hidden activation of the learned closure (default silu)"}}},
        {"Diagnostics and limits",
         {{"--diagnostic-interval N", "frames between summary rows"},
          {"--histogram-interval N", "frames between momentum histograms (0: last frame only)"},
          {"--particle-snapshot-interval N", "frames between snapshots (0: none)"},
          {"--walltime-hours X", "stop after X hours (0: no limit)"},
          {"--walltime-seconds X", "stop after X seconds (0: no limit)"},
          {"--walltime-reserve-minutes X", "stop X minutes before the limit"},
          {"--walltime-reserve-seconds X", "stop X seconds before the limit"}}},
    };
    std::cout << "kpt_run: reconnection calibration loop on the GPU push\n";
    for (const auto& g : groups) {
        std::cout << "\n" << g.title << ":\n";
        for (const HelpEntry& e : g.entries)
            std::cout << "  " << std::left << std::setw(36) << e.option << e.text << "\n";
    }
}

// parse_settings (src/main.cpp:2444-2611)
RunSettings parse(int argc, char** argv) {
    RunSettings s;
    if (kpt_default_host_settings(&s.host) != KPT_OK) throw std::runtime_error("host settings");
    std::string profile = s.profile_name;
    for (int i = 1; i < argc; ++i)
        if (std::string(argv[i]) == "--profile") {
            if (i + 1 >= argc) throw std::runtime_error("Missing value for --profile.");
            profile = argv[++i];
        }
    apply_profile(s, profile);
    bool start_set = false;
    int frames = -1;
    kpt_host_settings& h = s.host;
    for (int i = 1; i < argc; ++i) {
        const std::string a(argv[i]);
        auto val = [&]() { return next_value(argc, argv, i, a); };
        if (a == "--help" || a == "-h") {
            print_help();
            std::exit(0);
        } else if (a == "--profile") {
            (void)val();
        } else if (a == "--transport") {
            const std::string m = val();
            if (m == "parker") s.transport = Transport::Parker;
            else if (m == "focused") s.transport = Transport::Focused;
            else throw std::runtime_error("Unknown transport model: " + m);
        } else if (a == "--focused-transport") {
            s.transport = Transport::Focused;
        } else if (a == "--parker-transport") {
            s.transport = Transport::Parker;
        } else if (a == "--restart-particle-snapshot") {
            s.restart_particle_snapshot_path = val();
        } else if (a == "--overwrite-output") {
            s.overwrite_output = true;
        } else if (a == "--field-dir") {
            s.field_dir = val();
        } else if (a == "--output-dir") {
            s.output_dir = val();
        } else if (a == "--start-frame") {
            s.start_frame = std::stoi(val());
            start_set = true;
            if (frames > 0) s.end_frame = s.start_frame + frames;
        } else if (a == "--end-frame") {
            s.end_frame = std::stoi(val());
        } else if (a == "--frames") {
            frames = std::stoi(val());
            s.end_frame = s.start_frame + frames;
        } else if (a == "--particles-per-frame") {
            s.particles_per_frame = std::stoull(val());
        } else if (a == "--rank-scale") {
            s.rank_scale = std::stoull(val());
        } else if (a == "--capacity") {
            s.particle_capacity = std::stoull(val());
        } else if (a == "--seed") {
            s.seed = std::stoull(val());
        } else if (a == "--dt-out") {
            h.mhd_dt = std::stod(val());
        } else if (a == "--diagnostic-interval") {
            s.diagnostic_interval = std::stoi(val());
        } else if (a == "--histogram-interval") {
            s.histogram_interval = std::stoi(val());
        } else if (a == "--particle-snapshot-interval") {
            s.particle_snapshot_interval = std::stoi(val());
        } else if (a == "--max-steps-per-interval") {
            s.max_particle_steps_per_interval = std::stoi(val());
        } else if (a == "--walltime-hours") {
            s.walltime_limit_seconds = 3600.0 * std::stod(val());
        } else if (a == "--walltime-seconds") {
            s.walltime_limit_seconds = std::stod(val());
        } else if (a == "--walltime-reserve-minutes") {
            s.walltime_reserve_seconds = 60.0 * std::stod(val());
        } else if (a == "--walltime-reserve-seconds") {
            s.walltime_reserve_seconds = std::stod(val());
        } else if (a == "--particle-v0") {
            h.particle_v0 = std::stod(val());
        } else if (a == "--duu0") {
            h.duu0 = std::stod(val());
        } else if (a == "--split-ratio") {
            h.split_ratio = std::stod(val());
        } else if (a == "--pmin-split-over-p0") {
            h.pmin_split_over_p0 = std::stod(val());
        } else if (a == "--no-split") {
            s.split_particles = false;
        } else if (a == "--no-time-interp") {
            s.time_interpolation = false;
        } else if (a == "--resumable-snapshots") {
            s.resumable_snapshots = true;
        } else if (a == "--allow-stretched-fields") {
            s.allow_stretched_fields = true;
        } else if (a == "--synchronous-field-reads") {
            s.stream_fields = false;
        } else if (a == "--precision") {
            const std::string p = val();
            if (p == "fp64") s.precision = KPT_PREC_FP64;
            else if (p == "fp32") s.precision = KPT_PREC_FP32;
            else throw std::runtime_error("Unknown precision: " + p);
        } else if (a == "--push-chunks") {
            s.push_chunks = std::stoi(val());
        } else if (a == "--closure") {
            const std::string c = val();
            if (c == "reduced") s.closure = KPT_CLOSURE_REDUCED;
            else if (c == "qlt") s.closure = KPT_CLOSURE_QLT;
            else if (c == "mlp") s.closure = KPT_CLOSURE_MLP;
            else throw std::runtime_error("Unknown pitch-angle closure: " + c);
        } else if (a == "--qlt-sigma2") {
            s.qlt_sigma2 = std::stod(val());
        } else if (a == "--qlt-correlation-length") {
            s.qlt_correlation_length = std::stod(val());
        } else if (a == "--qlt-spectral-index") {
            s.qlt_spectral_index = std::stod(val());
        } else if (a == "--mlp-train-steps") {
            s.mlp_train_steps = static_cast<uint32_t>(std::stoul(val()));
        } else if (a == "--mlp-activation") {
            const std::string c = val();
            if (c == "silu") s.mlp_activation = KPT_MLP_SILU;
            else if (c == "hard-swish") s.mlp_activation = KPT_MLP_HARD_SWISH;
            else throw std::runtime_error("Unknown network activation: " + c);
        } else if (a == "--device") {
            s.device = std::stoi(val());
        } else {
            throw std::runtime_error("Unknown option: " + a);
        }
    }
    if (s.restart_enabled()) {
        const int rf = snapshot_frame(s.restart_particle_snapshot_path);
        if (rf < 0)
            throw std::runtime_error(
                "Cannot infer restart frame from particle snapshot name; expected particles_XXXXX.bin.");
        if (start_set && s.start_frame != rf)
            throw std::runtime_error("Restart particle snapshot frame does not match --start-frame.");
        if (!start_set) {
            s.start_frame = rf;
            if (frames > 0) s.end_frame = s.start_frame + frames;
        }
    }
    h.max_particle_steps_per_interval = s.max_particle_steps_per_interval;
    h.split_particles = s.split_particles ? 1 : 0;
    h.time_interpolation = s.time_interpolation ? 1 : 0;
    h.focused_transport = s.transport == Transport::Focused ? 1 : 0;
    s.validate();
    return s;
}

// require_fresh_output_directory (src/main.cpp:475-488)
void require_fresh_output(const RunSettings& s) {
    if (s.restart_enabled() || s.overwrite_output) return;
    const std::filesystem::path dir(s.output_dir);
    if (!std::filesystem::exists(dir)) return;
    if (!std::filesystem::is_directory(dir))
        throw std::runtime_error("Reconnection output path exists but is not a directory.");
    for (const auto& e : std::filesystem::directory_iterator(dir)) {
        const std::string n = e.path().filename().string();
        auto has = [&](const std::string& pre, const std::string& suf) {
            return n.size() > pre.size() + suf.size() && n.compare(0, pre.size(), pre) == 0 &&
                   n.compare(n.size() - suf.size(), suf.size(), suf) == 0;
        };
        if (n == "summary.csv" || n == "run.log" || has("particles_", ".bin") || has("momentum_histogram_", ".csv"))
            throw std::runtime_error(
                "Output directory already contains reconnection solver outputs. Use --restart-particle-snapshot "
                "to continue a completed-frame snapshot, choose a new --output-dir, or pass --overwrite-output to "
                "replace the existing run.");
    }
}

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

struct Ctx {
    kpt_ctx* c = nullptr;
    ~Ctx() {
        if (c) kpt_destroy(c);
    }
    void check(kpt_status st, const char* what) const {
        if (st != KPT_OK)
            throw std::runtime_error(std::string(what) + ": " + (c ? kpt_last_error(c) : kpt_last_create_error()));
    }
};

struct HostField {
    int32_t nx = 0, ny = 0;
    std::vector<double> xe, ye, planes[6];
};

HostField read_field(const std::string& path, bool allow_stretched) {
    HostField f;
    char err[512] = {0};
    const int32_t nonuniform = allow_stretched ? 1 : 0;
    if (kpt_compact_field_read(path.c_str(), &f.nx, &f.ny, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr, nonuniform, err, sizeof(err)) != KPT_OK)
        throw std::runtime_error(err);
    f.xe.resize(f.nx + 1);
    f.ye.resize(f.ny + 1);
    for (auto& p : f.planes) p.resize(static_cast<size_t>(f.nx) * f.ny);
    if (kpt_compact_field_read(path.c_str(), &f.nx, &f.ny, f.xe.data(), f.ye.data(), f.planes[0].data(),
                               f.planes[1].data(), f.planes[2].data(), f.planes[3].data(), f.planes[4].data(),
                               f.planes[5].data(), nonuniform, err, sizeof(err)) != KPT_OK)
        throw std::runtime_error(err);
    return f;
}

void upload(Ctx& ctx, int slot, int model, const HostField& f) {
    ctx.check(kpt_field_upload(ctx.c, slot, model, f.nx, f.ny, f.xe.data(), f.ye.data(), f.planes[0].data(),
                               f.planes[1].data(), f.planes[2].data(), f.planes[3].data(), f.planes[4].data(),
                               f.planes[5].data()),
              "field upload");
}

struct Log {
    std::ofstream f;
    void block(const std::string& b) {
        std::cout << b << std::flush;  // progress stays visible when stdout is a file
        if (f) {
            f << b;
            f.flush();
        }
    }
    void line(const std::string& l) { block(l + "\n"); }
};

int run(const RunSettings& rs) {
    require_fresh_output(rs);
    std::filesystem::create_directories(rs.output_dir);
    const std::string summary_path = rs.output_dir + "/summary.csv";
    const bool append = rs.restart_enabled() && std::filesystem::exists(summary_path);
    const bool has_content = append && std::filesystem::file_size(summary_path) > 0;
    std::ofstream summary(summary_path, append ? (std::ios::out | std::ios::app) : std::ios::out);
    if (!summary) throw std::runtime_error("Failed to open summary output.");
    char buf[4096];
    if (!has_content) {
        if (kpt_format_summary_header(buf, sizeof(buf)) != KPT_OK) throw std::runtime_error("summary header");
        summary << buf;
    }
    Log log;
    log.f.open(rs.output_dir + "/run.log", rs.restart_enabled() ? (std::ios::out | std::ios::app) : std::ios::out);
    if (!log.f) throw std::runtime_error("Failed to open run log output.");

    // A restart reseeds like the reference (src/main.cpp:2644-2650), except with
    // resumable snapshots: their particles keep their ids, and the streams keyed
    // by (seed, id, frame, sub-step) then continue the interrupted run.
    const uint64_t seed = rs.restart_enabled() && !rs.resumable_snapshots
                              ? rs.seed + 0x9e3779b97f4a7c15ull * static_cast<uint64_t>(rs.start_frame + 1)
                              : rs.seed;
    Ctx ctx;
    ctx.check(kpt_create(&ctx.c, rs.device, rs.particle_capacity, seed), "kpt_create");
    const int model = rs.transport == Transport::Parker ? KPT_MODEL_PARKER16 : KPT_MODEL_FOCUSED21;

    auto read_start = Clock::now();
    const bool stretched = rs.allow_stretched_fields;
    HostField lower = read_field(frame_path(rs.field_dir, "field", rs.start_frame, ".bin"), stretched);
    HostField upper = read_field(frame_path(rs.field_dir, "field", rs.start_frame + 1, ".bin"), stretched);
    upload(ctx, KPT_SLOT_LOWER, model, lower);
    upload(ctx, KPT_SLOT_UPPER, model, upper);
    const double initial_read_seconds = since(read_start);

    kpt_device_settings s;
    ctx.check(kpt_make_device_settings(&rs.host, lower.nx, lower.ny, lower.xe.data(), lower.ye.data(), &s),
              "device settings");
    s.precision = rs.precision;
    s.push_chunks = rs.push_chunks;
    s.closure = rs.closure;
    if (rs.closure != KPT_CLOSURE_REDUCED) {
        s.qlt_sigma2 = rs.qlt_sigma2;
        s.qlt_correlation_length = rs.qlt_correlation_length;
        s.qlt_spectral_index = rs.qlt_spectral_index;
    }
    if (rs.restart_enabled()) {
        // The particles of the closed interval start_frame - 1 carry its end time
        // t0 + mhd_dt exactly (the last sub-step is truncated to it); the restart
        // time is that value, which differs from start_frame * mhd_dt by at most
        // one rounding.
        const double restart_time =
            rs.start_frame > 0 ? static_cast<double>(rs.start_frame - 1) * rs.host.mhd_dt + rs.host.mhd_dt : 0.0;
        ctx.check(kpt_snapshot_read(ctx.c, rs.restart_particle_snapshot_path.c_str(), rs.host.charge,
                                    rs.host.split_ratio, restart_time, s.dt_min),
                  "restart snapshot");
        uint64_t n = 0;
        kpt_particles_count(ctx.c, &n);
        std::ostringstream l;
        l << "Restarted from particle snapshot " << rs.restart_particle_snapshot_path << " at frame "
          << rs.start_frame << " loaded_particles=" << n << " configured_capacity=" << rs.particle_capacity
          << " effective_seed=" << seed << " (particle streams are keyed by seed, frame and particle id)";
        log.line(l.str());
    }
    if (rs.closure == KPT_CLOSURE_MLP) {
        kpt_mlp_train_opts o;
        kpt_mlp_default_train_opts(&o);
        o.steps = rs.mlp_train_steps;
        o.hull.activation = rs.mlp_activation;
        kpt_mlp_report r;
        ctx.check(kpt_mlp_train(ctx.c, &o, &r), "pitch-angle network training");
        std::ostringstream l;
        l << "Trained the D_mumu network: steps=" << r.steps << " train_ms=" << r.train_ms
          << " rms(ln R)=" << r.val_rms_lnr << " rms(dD/D)=" << r.val_rms_rel_d;
        log.line(l.str());
    }

    std::ostringstream header;
    header << "Reconnection 2D calibration\n"
           << "  profile=" << rs.profile_name << '\n'
           << "  transport=" << (rs.transport == Transport::Parker ? "parker" : "focused") << '\n'
           << "  field_dir=" << rs.field_dir << '\n'
           << "  output_dir=" << rs.output_dir << '\n'
           << "  restart_particle_snapshot="
           << (rs.restart_enabled() ? rs.restart_particle_snapshot_path : std::string("disabled")) << '\n'
           << "  overwrite_output=" << (rs.overwrite_output ? "enabled" : "disabled") << '\n'
           << "  frames=" << rs.start_frame << ".." << rs.end_frame << " (" << rs.interval_count() << " intervals)\n"
           << "  particles_per_frame=" << rs.particles_per_frame * rs.rank_scale << " (" << rs.particles_per_frame
           << " * rank_scale " << rs.rank_scale << ")\n"
           << "  capacity=" << rs.particle_capacity << '\n'
           << "  dt_out=" << rs.host.mhd_dt << '\n'
           << "  histogram_interval=" << rs.histogram_interval << '\n'
           << "  particle_snapshot_interval=" << rs.particle_snapshot_interval << '\n'
           << "  walltime_limit_s=";
    if (rs.walltime_limit_enabled())
        header << rs.walltime_limit_seconds << " walltime_reserve_s=" << rs.walltime_reserve_seconds
               << " walltime_stop_s=" << rs.walltime_stop_seconds() << '\n';
    else
        header << "disabled\n";
    header << "  kpara0=" << rs.host.kpara0 << " kperp/kpara=" << rs.host.kperp_over_kpara << '\n'
           << "  particle_v0=" << rs.host.particle_v0 << " duu0=" << rs.host.duu0 << '\n'
           << "  device=" << kpt_version() << " precision=" << (rs.precision == KPT_PREC_FP32 ? "fp32" : "fp64")
           << " push_chunks=" << rs.push_chunks << '\n'
           << "  field_reads=" << (rs.stream_fields ? "staged during the push" : "synchronous") << '\n';
    log.block(header.str());

    // The frame after the upper one loads in the background during each push.
    const auto stage = [&](int f) {
        if (rs.stream_fields && f <= rs.end_frame)
            ctx.check(kpt_field_stage_compact(ctx.c, frame_path(rs.field_dir, "field", f, ".bin").c_str(), model,
                                              stretched ? 1 : 0),
                      "field stage");
    };
    stage(rs.start_frame + 2);
    const auto run_start = Clock::now();
    bool stopped = false;
    int last = rs.start_frame;
    std::vector<double> hist(128);
    for (int frame = rs.start_frame; frame < rs.end_frame; ++frame) {
        const auto frame_start = Clock::now();
        kpt_frame_timing timing{};
        timing.read_seconds = frame == rs.start_frame ? initial_read_seconds : 0.0;
        if (frame > rs.start_frame) {
            const auto t = Clock::now();
            if (rs.stream_fields) {
                ctx.check(kpt_field_advance_staged(ctx.c, nullptr), "field advance");
                stage(frame + 2);
            } else {
                ctx.check(kpt_field_advance(ctx.c), "field advance");
                upload(ctx, KPT_SLOT_UPPER, model,
                       read_field(frame_path(rs.field_dir, "field", frame + 1, ".bin"), stretched));
            }
            timing.read_seconds = since(t);
        }
        const double t0 = static_cast<double>(frame) * rs.host.mhd_dt;
        auto t = Clock::now();
        uint64_t injected = 0;
        ctx.check(kpt_inject(ctx.c, &s, nullptr, rs.particles_per_frame * rs.rank_scale, t0,
                             static_cast<uint32_t>(frame), &injected),
                  "inject");
        timing.injected = injected;
        timing.inject_seconds = since(t);
        t = Clock::now();
        kpt_push_stats ps;
        ctx.check(kpt_push_interval(ctx.c, &s, t0, static_cast<uint32_t>(frame), &ps), "push");
        timing.move_seconds = since(t);
        t = Clock::now();
        ctx.check(kpt_split(ctx.c, &s, &timing.split, &timing.skipped_split_capacity), "split");
        timing.split_seconds = since(t);

        const double frame_s = since(frame_start);
        const double elapsed_s = since(run_start);
        const bool walltime = rs.walltime_limit_enabled() && elapsed_s >= rs.walltime_stop_seconds();
        const bool finishing = frame + 1 == rs.end_frame || walltime;
        if ((frame + 1 - rs.start_frame) % rs.diagnostic_interval == 0 || finishing) {
            kpt_summary sm;
            ctx.check(kpt_summary_compute(ctx.c, &sm), "summary");
            if (kpt_format_summary_row(buf, sizeof(buf), frame + 1, (frame + 1) * rs.host.mhd_dt, &sm, &timing,
                                       frame_s, elapsed_s) != KPT_OK)
                throw std::runtime_error("summary row");
            summary << buf;
            summary.flush();
            std::ostringstream l;
            l << "frame " << frame + 1 << " active=" << sm.active << " stored=" << sm.stored
              << " pmax=" << std::setprecision(6) << sm.p_max << " read_s=" << timing.read_seconds
              << " inject_s=" << timing.inject_seconds << " move_s=" << timing.move_seconds
              << " split_s=" << timing.split_seconds << " frame_s=" << frame_s << " elapsed_s=" << elapsed_s;
            log.line(l.str());
        }
        if ((rs.histogram_interval > 0 && (frame + 1 - rs.start_frame) % rs.histogram_interval == 0) || finishing) {
            ctx.check(kpt_histogram_logp(ctx.c, 128, rs.host.pmin, rs.host.pmax, hist.data()), "histogram");
            if (kpt_write_histogram_csv(frame_path(rs.output_dir, "momentum_histogram_", frame + 1, ".csv").c_str(),
                                        128, rs.host.pmin, rs.host.pmax, hist.data()) != KPT_OK)
                throw std::runtime_error("Failed to open momentum histogram output.");
        }
        if (rs.particle_snapshot_interval > 0 &&
            ((frame + 1 - rs.start_frame) % rs.particle_snapshot_interval == 0 || finishing)) {
            const auto ts = Clock::now();
            const std::string path = frame_path(rs.output_dir, "particles_", frame + 1, ".bin");
            ctx.check(rs.resumable_snapshots
                          ? kpt_snapshot_write_resumable(ctx.c, path.c_str(), rs.host.charge, rs.host.split_ratio,
                                                         rs.host.p0)
                          : kpt_snapshot_write(ctx.c, path.c_str(), rs.host.charge, rs.host.split_ratio, rs.host.p0),
                      "snapshot");
            std::ostringstream l;
            l << "wrote particle snapshot " << path << " in " << since(ts) << " s";
            log.line(l.str());
        }
        last = frame + 1;
        if (walltime) {
            stopped = true;
            std::ostringstream l;
            l << "Reached walltime stop after frame " << last << " elapsed_s=" << elapsed_s
              << " stop_s=" << rs.walltime_stop_seconds() << " limit_s=" << rs.walltime_limit_seconds
              << " reserve_s=" << rs.walltime_reserve_seconds;
            log.line(l.str());
            break;
        }
    }
    std::ostringstream footer;
    footer << (stopped ? "Stopped by walltime after frame " + std::to_string(last) + " in " : std::string("Finished in "))
           << since(run_start) << " s. Summary: " << rs.output_dir << "/summary.csv\n";
    log.block(footer.str());
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const RunSettings s = parse(argc, argv);
        return run(s);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << '\n';
        return 1;
    }
}
