// Ptarmigan CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// LMA fields (src/field/plane_wave.rs, src/field/focused_laser.rs) -- placeholder.
#include <cstdio>
#include <cstdlib>
#include "oracle.hpp"

namespace orc {
struct RadEvent;
struct PairEvent;
[[noreturn]] static void todo(const char* w) { std::fprintf(stderr, "oracle: %s not restated yet\n", w); std::abort(); }
double lma_max_timestep(const Laser&) { todo("lma_max_timestep"); }
bool lma_contains(const Laser&, V4) { todo("lma_contains"); }
double lma_ideal_initial_z(const Laser&) { todo("lma_ideal_initial_z"); }
PushOut lma_push(const Laser&, V4, V4, double, double, Eqn) { todo("lma_push"); }
bool lma_radiate(const Laser&, V4, V4, double, Xoshiro&, bool, RadEvent*) { todo("lma_radiate"); }
double lma_pair_create(const Laser&, V4, V4, SV, double, Xoshiro&, double, SV*, bool*, PairEvent*) { todo("lma_pair_create"); }
}  // namespace orc
