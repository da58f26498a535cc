// kpt_convert_mhd — converts Fortran MHD output frames into the compact-field
// files kpt_run reads (apps/fortran_mhd_to_compact_field.cpp:171-250). Host C++;
// the file layouts are the reference's, the code is this build's own.
//
// Input (in --input-dir):
//   mhd_config.dat   13 f64 (dx, dy, dz, xmin, ymin, zmin, xmax, ymax, zmax, lx,
//                    ly, lz, dt_out) then 13 i32 (nx, ny, nz, nxs, nys, nzs, topox,
//                    topoy, topoz, nvar, bcx, bcy, bcz); 2D only (nz = 1)
//   mhd_data_NNNN    f32 [nz][ny + 4][nx + 4][8], variables vx, vy, vz, -, bx, by,
//                    bz, - per cell, two ghost layers in x and y
// Output (in --output-dir): fieldNNNNN.bin, the compact field of
// include/AthenaFieldReader.hpp:274-288 (i32 nx, ny; f64 x and y faces
// xmin + i (xmax - xmin) / nx; f64 planes Bx, By, Bz, Vx, Vy, Vz of the physical
// cells, x fastest).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct Config {
    double real[13];
    int32_t integer[13];
    double xmin() const { return real[3]; }
    double ymin() const { return real[4]; }
    double xmax() const { return real[6]; }
    double ymax() const { return real[7]; }
    int32_t nx() const { return integer[0]; }
    int32_t ny() const { return integer[1]; }
    int32_t nz() const { return integer[2]; }
};

struct Options {
    std::filesystem::path input_dir, output_dir;
    std::string config_name = "mhd_config.dat";
    int start_frame = 0, end_frame = 200;
};

std::string numbered(const char* prefix, int width, int frame, const char* suffix) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%s%0*d%s", prefix, width, frame, suffix);
    return buf;
}

Config read_config(const std::filesystem::path& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("failed to open MHD config: " + path.string());
    Config c{};
    in.read(reinterpret_cast<char*>(c.real), sizeof(c.real));
    in.read(reinterpret_cast<char*>(c.integer), sizeof(c.integer));
    if (!in) throw std::runtime_error("failed to read MHD config payload: " + path.string());
    if (c.nx() <= 0 || c.ny() <= 0 || c.nz() != 1)
        throw std::runtime_error("only positive 2D MHD configurations are supported");
    return c;
}

void put(std::ofstream& out, const void* p, size_t bytes) {
    out.write(static_cast<const char*>(p), static_cast<std::streamsize>(bytes));
    if (!out) throw std::runtime_error("failed while writing double payload");
}

void convert(const Config& c, const std::filesystem::path& in_path, const std::filesystem::path& out_path) {
    constexpr int kVars = 8, kGhost = 2;
    const int nx = c.nx(), ny = c.ny();
    const size_t sx = static_cast<size_t>(nx) + 2 * kGhost, sy = static_cast<size_t>(ny) + 2 * kGhost;
    std::vector<float> raw(kVars * sx * sy);
    std::ifstream in(in_path, std::ios::binary);
    if (!in) throw std::runtime_error("failed to open MHD frame: " + in_path.string());
    in.read(reinterpret_cast<char*>(raw.data()), static_cast<std::streamsize>(raw.size() * sizeof(float)));
    if (!in) throw std::runtime_error("failed to read MHD frame: " + in_path.string());

    std::ofstream out(out_path, std::ios::binary);
    if (!out) throw std::runtime_error("failed to open compact frame: " + out_path.string());
    const int32_t dims[2] = {nx, ny};
    put(out, dims, sizeof(dims));
    const double lo[2] = {c.xmin(), c.ymin()}, hi[2] = {c.xmax(), c.ymax()};
    for (int d = 0; d < 2; ++d) {
        const int n = dims[d];
        const double step = (hi[d] - lo[d]) / static_cast<double>(n);
        std::vector<double> faces(static_cast<size_t>(n) + 1);
        for (int k = 0; k <= n; ++k) faces[k] = lo[d] + static_cast<double>(k) * step;
        put(out, faces.data(), faces.size() * sizeof(double));
    }
    // compact order Bx, By, Bz, Vx, Vy, Vz from the MHD variables 4, 5, 6, 0, 1, 2
    static const int kVar[6] = {4, 5, 6, 0, 1, 2};
    std::vector<double> plane(static_cast<size_t>(nx) * ny);
    for (int v : kVar) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                plane[static_cast<size_t>(j) * nx + i] =
                    raw[v + kVars * ((static_cast<size_t>(i) + kGhost) + sx * (static_cast<size_t>(j) + kGhost))];
        put(out, plane.data(), plane.size() * sizeof(double));
    }
}

Options parse(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&](const char* opt) -> std::string {
            if (i + 1 >= argc) throw std::runtime_error(std::string("missing value for ") + opt);
            return argv[++i];
        };
        if (a == "--input-dir") o.input_dir = value("--input-dir");
        else if (a == "--output-dir") o.output_dir = value("--output-dir");
        else if (a == "--config") o.config_name = value("--config");
        else if (a == "--start-frame") o.start_frame = std::stoi(value("--start-frame"));
        else if (a == "--end-frame") o.end_frame = std::stoi(value("--end-frame"));
        else if (a == "--help" || a == "-h") {
            std::cout << "Usage: kpt_convert_mhd --input-dir PATH --output-dir PATH [--config NAME] "
                         "[--start-frame N] [--end-frame N]\n";
            std::exit(0);
        } else {
            throw std::runtime_error("unknown option: " + a);
        }
    }
    if (o.input_dir.empty() || o.output_dir.empty())
        throw std::runtime_error("--input-dir and --output-dir are required");
    if (o.end_frame < o.start_frame) throw std::runtime_error("end frame must be greater than or equal to start frame");
    return o;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Options o = parse(argc, argv);
        std::filesystem::create_directories(o.output_dir);
        const Config c = read_config(o.input_dir / o.config_name);
        std::cout << "Converting Fortran MHD frames " << o.start_frame << ".." << o.end_frame << " to "
                  << o.output_dir << "\n  nx=" << c.nx() << " ny=" << c.ny() << " xmin=" << c.xmin()
                  << " xmax=" << c.xmax() << " ymin=" << c.ymin() << " ymax=" << c.ymax() << '\n';
        for (int f = o.start_frame; f <= o.end_frame; ++f) {
            const auto out = o.output_dir / numbered("field", 5, f, ".bin");
            convert(c, o.input_dir / numbered("mhd_data_", 4, f, ""), out);
            std::cout << "frame " << f << " -> " << out << '\n';
        }
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << '\n';
        return 1;
    }
    return 0;
}
