// Runs a few oracle steps (float and double, with neighbour recording and
// 4 OpenMP threads) on a jittered lattice with a wall hit, for a build under
// AddressSanitizer and UndefinedBehaviorSanitizer.  Test infrastructure only.
#include <cstdio>
#include <random>
#include <vector>

#include "sph_oracle.hpp"

template <class T>
static int run() {
    orc::Params p{};
    p.particle_radius = 0.005f;
    p.kernel_radius = 0.02f;
    p.rest_density = 1000.f;
    p.pressure_stiffness = 400.f;
    p.viscosity = 0.05f;
    p.boundary_restitution = 0.1f;
    p.attractive_pressure_ratio = 0.1f;
    p.gravity[1] = -9.81f;
    p.dt = 1e-4f;
    for (int k = 0; k < 3; ++k) { p.box_min[k] = 0.f; p.box_max[k] = 0.5f; }
    p.x_subdivision = 4;
    p.cfl_number = 0.3f;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(-3e-4, 3e-4);
    std::vector<T> pos, vel;
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 10; ++j)
            for (int k = 0; k < 8; ++k) {
                pos.push_back(T(0.005 + 0.01 * i + u(rng)));
                pos.push_back(T(0.005 + 0.01 * j + u(rng)));
                pos.push_back(T(0.005 + 0.01 * k + u(rng)));
                vel.push_back(T(0));
                vel.push_back(T(-0.5));
                vel.push_back(T(0));
            }
    orc::Sim<T> s;
    s.set_params(p);
    s.nthreads = 4;
    s.record_nbrs = true;
    s.set_state((uint32_t)(pos.size() / 3), pos.data(), vel.data(), nullptr);
    for (int t = 0; t < 5; ++t) s.step();
    for (uint32_t i = 0; i < s.n; ++i)
        if (!(s.y[i] >= T(0.005) - T(1e-6))) return 1;
    return 0;
}

int main() {
    const int rc = run<float>() | run<double>();
    std::printf("oracle sanitizer run rc=%d\n", rc);
    return rc;
}
