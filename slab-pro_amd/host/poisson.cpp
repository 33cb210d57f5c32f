// poisson.cpp -- Bridson Poisson-disk sampling (see poisson.hpp).
#include "poisson.hpp"

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>

namespace slabhost {

namespace {

// glibc rand() sequence of a fresh process, from a private state.
class FreshRand {
  public:
    FreshRand() {
        if (initstate_r(1u, state_, sizeof(state_), &data_) != 0) throw std::runtime_error("initstate_r failed");
    }
    int next() {
        int32_t r = 0;
        random_r(&data_, &r);
        return (int)r;
    }
    // U in [0, 1] as rand() / RAND_MAX
    double unit() { return (double)next() / (double)RAND_MAX; }

  private:
    random_data data_{};
    char state_[128] = {};
};

} // namespace

std::vector<std::pair<double, double>> poisson_disk(double dx, double Lx, double Ly, int ppc) {
    const double r = std::sqrt(15.0 / 24.0 * dx * dx / ppc);
    const double r2 = r * r;
    const double a = r / std::sqrt(2);
    const int ncx = (int)(Lx / a) + 1, ncy = (int)(Ly / a) + 1;
    std::vector<int> cell((size_t)ncx * ncy, -1);
    std::vector<std::pair<double, double>> s;
    std::vector<int> active;
    FreshRand rng;

    auto cell_of = [&](double x, double y, int& cx, int& cy) {
        cx = (int)std::floor(x / a);
        cy = (int)std::floor(y / a);
    };
    // a sample is accepted when no earlier sample lies closer than r; within
    // cells of side r/sqrt(2) those can only sit in the 5x5 block around the
    // candidate's cell without its four corners
    auto far_enough = [&](double x, double y) {
        int cx, cy;
        cell_of(x, y, cx, cy);
        for (int di = -2; di <= 2; ++di) {
            for (int dj = -2; dj <= 2; ++dj) {
                if ((di == -2 || di == 2) && (dj == -2 || dj == 2)) continue;
                const int i = cx + di, j = cy + dj;
                if (i < 0 || i >= ncx || j < 0 || j >= ncy) continue;
                const int q = cell[(size_t)i * ncy + j];
                if (q < 0) continue;
                const double ex = s[q].first - x, ey = s[q].second - y;
                if (ex * ex + ey * ey < r2) return false;
            }
        }
        return true;
    };

    const double x0 = rng.unit() * Lx;
    const double y0 = rng.unit() * Ly;
    s.emplace_back(x0, y0);
    {
        int cx, cy;
        cell_of(x0, y0, cx, cy);
        cell[(size_t)cx * ncy + cy] = 0;
    }
    active.push_back(0);
    while (!active.empty()) {
        const int pick = rng.next() % (int)active.size();
        const double rx = s[active[pick]].first, ry = s[active[pick]].second;
        bool found = false;
        double nx = 0, ny = 0;
        for (int attempt = 0; attempt < 30; ++attempt) {
            const double rho = rng.unit() * r + r;
            const double theta = rng.unit() * 2 * M_PI;
            double sn, cs;
            sincos(theta, &sn, &cs); // one call for the pair, as a GCC build of the reference makes
            nx = rx + rho * cs;
            ny = ry + rho * sn;
            if (nx >= 0 && nx <= Lx && ny >= 0 && ny <= Ly && far_enough(nx, ny)) {
                found = true;
                break;
            }
        }
        if (!found) {
            active.erase(active.begin() + pick);
            continue;
        }
        const int id = (int)s.size();
        s.emplace_back(nx, ny);
        active.push_back(id);
        int cx, cy;
        cell_of(nx, ny, cx, cy);
        cell[(size_t)cx * ncy + cy] = id;
    }
    return s;
}

} // namespace slabhost
