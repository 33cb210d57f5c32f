// poisson.hpp -- Poisson-disk sampling of a rectangle (Bridson's algorithm,
// "Fast Poisson disk sampling in arbitrary dimensions", SIGGRAPH 2007) with
// the parameters and random-number consumption of src/poisson_disk_sampling.hpp:
//   r = sqrt(15/24 * dx^2 / ppc), background cells of side r/sqrt(2),
//   k = 30 candidates per active sample, annulus radius U*r + r,
//   angle U*2*pi, active sample picked as rand() % |active|,
//   exhausted samples removed from the active list in place (order kept).
// The random stream is glibc's rand() of a freshly started process
// (TYPE_3 additive generator seeded with 1), drawn from a private state so
// that the caller's rand() state is untouched.
#pragma once
#include <utility>
#include <vector>

namespace slabhost {

std::vector<std::pair<double, double>> poisson_disk(double dx, double Lx, double Ly, int ppc);

} // namespace slabhost
