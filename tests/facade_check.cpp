// facade_check.cpp — drives the C++ facade (include/kpt.hpp) the way code written
// against the reference's generic API would: a ParticleSystem, a ParkerSolver or
// FocusTransportSolver with set_coefficient_fields, injected-noise
// advance_stochastic steps, one advance_deterministic step, and a
// write_binary_snapshot / read_binary_snapshot round trip. Inputs and outputs are
// raw little-endian files written and read by tests/test_gpu_facade.py, which
// compares the outputs with the CPU oracle.
//
// usage: facade_check <input> <output> <snapshot path>
// input:  int32 solver (0 Parker, 1 focused), nx, ny, n, steps;
//         the solver's parameter struct (raw bytes);
//         f64 x_edges[nx+1], y_edges[ny+1], planes[6][ny][nx];
//         f64 x[n], y[n], p[n], mu[n], weight[n]; int32 status[n]; u64 id[n];
//         f64 noise[n][steps][4]
// output: f64 dt[steps + 1]; f64 x, y, p, mu, prev_x, prev_y, prev_step_x,
//         prev_step_y [n each]; int32 status[n]; int32 snapshot_round_trip_ok
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "../include/kpt.hpp"

namespace {

template <typename T>
void read_into(std::ifstream& in, T* dst, size_t n) {
    in.read(reinterpret_cast<char*>(dst), static_cast<std::streamsize>(n * sizeof(T)));
    if (!in) throw std::runtime_error("facade_check: short input file");
}

template <typename T>
void write_from(std::ofstream& out, const T* src, size_t n) {
    out.write(reinterpret_cast<const char*>(src), static_cast<std::streamsize>(n * sizeof(T)));
}

struct HostSet {
    std::vector<double> x, y, p, t, step_dt, w, mu, e0, px, py, psx, psy;
    std::vector<int32_t> status, level, key;
    std::vector<uint64_t> id;
    explicit HostSet(size_t n)
        : x(n), y(n), p(n), t(n), step_dt(n), w(n), mu(n), e0(n), px(n), py(n), psx(n), psy(n), status(n),
          level(n), key(n), id(n) {}
    kpt_particles_host view() {
        return kpt_particles_host{x.data(),      y.data(),  p.data(),  t.data(),  step_dt.data(), w.data(),
                                  mu.data(),     status.data(), level.data(), id.data(), e0.data(), px.data(),
                                  py.data(),     psx.data(), psy.data(), key.data()};
    }
};

bool same_bits(const HostSet& a, const HostSet& b, size_t n) {
    auto eq = [n](const void* u, const void* v, size_t es) { return std::memcmp(u, v, n * es) == 0; };
    return eq(a.x.data(), b.x.data(), 8) && eq(a.y.data(), b.y.data(), 8) && eq(a.p.data(), b.p.data(), 8) &&
           eq(a.mu.data(), b.mu.data(), 8) && eq(a.w.data(), b.w.data(), 8) && eq(a.e0.data(), b.e0.data(), 8) &&
           eq(a.px.data(), b.px.data(), 8) && eq(a.py.data(), b.py.data(), 8) &&
           eq(a.psx.data(), b.psx.data(), 8) && eq(a.psy.data(), b.psy.data(), 8) &&
           eq(a.status.data(), b.status.data(), 4) && eq(a.level.data(), b.level.data(), 4) &&
           eq(a.key.data(), b.key.data(), 4) && eq(a.id.data(), b.id.data(), 8);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: facade_check <input> <output> <snapshot path>\n");
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        int32_t hdr[5];
        read_into(in, hdr, 5);
        const int solver = hdr[0], nx = hdr[1], ny = hdr[2];
        const size_t n = static_cast<size_t>(hdr[3]);
        const uint32_t steps = static_cast<uint32_t>(hdr[4]);
        kpt_generic_parker_params parker{};
        kpt_generic_focus_params focus{};
        if (solver == 0)
            read_into(in, reinterpret_cast<char*>(&parker), sizeof(parker));
        else
            read_into(in, reinterpret_cast<char*>(&focus), sizeof(focus));
        std::vector<double> xe(nx + 1), ye(ny + 1), planes(6 * static_cast<size_t>(nx) * ny);
        read_into(in, xe.data(), xe.size());
        read_into(in, ye.data(), ye.size());
        read_into(in, planes.data(), planes.size());
        HostSet h(n);
        read_into(in, h.x.data(), n);
        read_into(in, h.y.data(), n);
        read_into(in, h.p.data(), n);
        read_into(in, h.mu.data(), n);
        read_into(in, h.w.data(), n);
        read_into(in, h.status.data(), n);
        read_into(in, h.id.data(), n);
        std::vector<double> noise(n * steps * 4);
        read_into(in, noise.data(), noise.size());

        kpt::ParticleSystem particles(n);
        kpt_particles_host up{h.x.data(), h.y.data(), h.p.data(), h.t.data(), h.step_dt.data(), h.w.data(),
                              h.mu.data(), h.status.data(), h.level.data(), h.id.data()};
        particles.upload(up, n);  // previous positions start at the positions
        particles.set_injected_noise(noise.data(), n, steps);

        kpt::FieldPlanes f;
        f.nx = nx;
        f.ny = ny;
        f.x_edges = xe.data();
        f.y_edges = ye.data();
        const size_t plane = static_cast<size_t>(nx) * ny;
        for (int c = 0; c < 3; ++c) {
            f.b[c] = planes.data() + c * plane;
            f.v[c] = planes.data() + (3 + c) * plane;
        }
        std::vector<double> dts;
        if (solver == 0) {
            kpt::ParkerSolver s(particles, parker);
            s.set_coefficient_fields(f);
            for (uint32_t k = 0; k < steps; ++k) dts.push_back(s.advance_stochastic());
            dts.push_back(s.advance_deterministic());
        } else {
            kpt::FocusSolver s(particles, focus);
            s.set_coefficient_fields(f);
            s.set_maximum_pitch_angle_cosine(focus.maximum_pitch_angle_cosine);
            for (uint32_t k = 0; k < steps; ++k) dts.push_back(s.advance_stochastic());
            dts.push_back(s.advance_deterministic());
        }

        HostSet got(n);
        kpt_particles_host view = got.view();
        particles.download(view, n);

        // write_binary_snapshot -> read_binary_snapshot into a second system
        particles.write_binary_snapshot(argv[3]);
        kpt::ParticleSystem restored(n);
        restored.read_binary_snapshot(argv[3]);
        HostSet back(n);
        kpt_particles_host bview = back.view();
        restored.download(bview, restored.particle_count());
        // time and step_dt are not part of the ParticleSystem record
        back.t = got.t;
        back.step_dt = got.step_dt;
        const int32_t round_trip = restored.particle_count() == n && same_bits(got, back, n) ? 1 : 0;

        std::ofstream out(argv[2], std::ios::binary);
        write_from(out, dts.data(), dts.size());
        for (const auto* v : {&got.x, &got.y, &got.p, &got.mu, &got.px, &got.py, &got.psx, &got.psy})
            write_from(out, v->data(), n);
        write_from(out, got.status.data(), n);
        write_from(out, &round_trip, 1);
        if (!out) throw std::runtime_error("facade_check: failed to write the output file");
        std::printf("facade_check: solver %d, %zu particles, %u stochastic steps + 1 deterministic, "
                    "snapshot round trip %s\n",
                    solver, n, steps, round_trip ? "ok" : "FAILED");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
