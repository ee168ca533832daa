// sph_oracle.cpp -- TEST INFRASTRUCTURE ONLY: host restatement of the
// reference SPH2S step (see sph_oracle.h for the precision contract).
// Compiled with -ffp-contract=off so that the FP32 cell-coordinate expressions
// round exactly like the reference's separate IEEE operations.

#include "sph_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr double kPi = 3.14159265358979323846;  // Solver.h:11 (M_PI)

// computeCoordinatesOff (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:20-38):
// c = clamp((int)((x + off - o)/spacing), 0, D-1), all in FP32.
inline int coord_off(float x, float off, float o, float sp, int dim) {
    float t = x + off;
    t = t - o;
    t = t / sp;
    int c = static_cast<int>(t);
    return std::max(0, std::min(c, dim - 1));
}

// computeCoordinates (Solver.cu:40-56): (int)((x - o)/spacing), clamped.
inline int coord(float x, float o, float sp, int dim) {
    float t = x - o;
    t = t / sp;
    int c = static_cast<int>(t);
    return std::max(0, std::min(c, dim - 1));
}

// computeHash (Solver.cu:58-68): cx + Dx*(cy + Dy*cz).
inline uint32_t hash_of(const oracle_grid& g, int cx, int cy, int cz) {
    return static_cast<uint32_t>(cx + g.dims[0] * (cy + g.dims[1] * cz));
}

inline uint32_t hash_pos(const oracle_grid& g, float x, float y, float z) {
    return hash_of(g, coord(x, g.origin[0], g.spacing, g.dims[0]),
                   coord(y, g.origin[1], g.spacing, g.dims[1]),
                   coord(z, g.origin[2], g.spacing, g.dims[2]));
}

uint32_t num_cells(const oracle_grid& g) {
    return static_cast<uint32_t>(g.dims[0]) * static_cast<uint32_t>(g.dims[1]) *
           static_cast<uint32_t>(g.dims[2]);
}

// Kernel coefficients (Solver.cu:103-167): Poly6 315/(64 pi h^9),
// spiky gradient -45/(pi h^6)(h-r)^2 r_hat, viscosity Laplacian 45/(pi h^6)(h-r).
struct Coeffs {
    double c6, c45;
    explicit Coeffs(double h) {
        const double h3 = h * h * h;
        c6 = 315.0 / (64.0 * kPi * h3 * h3 * h3);
        c45 = 45.0 / (kPi * h3 * h3);
    }
};

// Particle state in sorted slot order, FP64.
struct State {
    std::vector<double> x, v;     // AoS 3n
    std::vector<uint32_t> id;     // persistent ids
};

struct Neighbourhood {
    std::vector<uint32_t> hash_sorted, perm, start, end;
};

// computeNeighborhoods (Solver.cu:1098-1140): hash -> stable sort ->
// cell start/end -> reorder.  Hashes are taken from the FP32 value of the
// position, as the reference state is FP32.
void build_neighbourhood(const oracle_params& p, State& s, Neighbourhood& nb) {
    const uint32_t n = static_cast<uint32_t>(s.id.size());
    std::vector<uint32_t> hash(n);
    for (uint32_t i = 0; i < n; ++i)
        hash[i] = hash_pos(p.grid, static_cast<float>(s.x[3 * i]),
                           static_cast<float>(s.x[3 * i + 1]),
                           static_cast<float>(s.x[3 * i + 2]));
    nb.hash_sorted.resize(n);
    nb.perm.resize(n);
    oracle_stable_sort(hash.data(), n, nb.hash_sorted.data(), nb.perm.data());
    const uint32_t C = num_cells(p.grid);
    nb.start.assign(C, kEmpty);
    nb.end.assign(C, kEmpty);
    oracle_cell_ranges(nb.hash_sorted.data(), n, C, nb.start.data(), nb.end.data());
    // reorderAndComputeCellStartEndD (Solver.cu:478-534): gather x, v by perm.
    State t;
    t.x.resize(3 * static_cast<size_t>(n));
    t.v.resize(3 * static_cast<size_t>(n));
    t.id.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t j = nb.perm[i];
        for (int d = 0; d < 3; ++d) {
            t.x[3 * i + d] = s.x[3 * j + d];
            t.v[3 * i + d] = s.v[3 * j + d];
        }
        t.id[i] = s.id[j];
    }
    s = std::move(t);
}

// The 27-cell (cs..ce inclusive) stencil of Solver.cu:559-593: bounds from the
// FP32 position offset by -h/+h.
template <class F>
void for_each_stencil_cell(const oracle_params& p, const double* xi, F&& f) {
    const oracle_grid& g = p.grid;
    const float x = static_cast<float>(xi[0]), y = static_cast<float>(xi[1]),
                z = static_cast<float>(xi[2]);
    const float h = p.h_f32;
    const int csx = coord_off(x, -h, g.origin[0], g.spacing, g.dims[0]);
    const int csy = coord_off(y, -h, g.origin[1], g.spacing, g.dims[1]);
    const int csz = coord_off(z, -h, g.origin[2], g.spacing, g.dims[2]);
    const int cex = coord_off(x, h, g.origin[0], g.spacing, g.dims[0]);
    const int cey = coord_off(y, h, g.origin[1], g.spacing, g.dims[1]);
    const int cez = coord_off(z, h, g.origin[2], g.spacing, g.dims[2]);
    for (int cz = csz; cz <= cez; ++cz)
        for (int cy = csy; cy <= cey; ++cy)
            for (int cx = csx; cx <= cex; ++cx) f(hash_of(g, cx, cy, cz));
}

// computeDensitiesPressuresD + computeDensityCell (Solver.cu:198-229,536-599):
// rho_i = m sum_{j: r < h} W_poly6(r), self included.
void densities(const oracle_params& p, const State& s, const Neighbourhood& nb,
               std::vector<double>& rho) {
    const uint32_t n = static_cast<uint32_t>(s.id.size());
    const Coeffs k(p.h);
    const double h2 = p.h * p.h;
    rho.assign(n, 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        const double* xi = &s.x[3 * i];
        double sum = 0.0;
        for_each_stencil_cell(p, xi, [&](uint32_t c) {
            const uint32_t b = nb.start[c], e = nb.end[c];
            if (b == kEmpty) return;
            for (uint32_t j = b; j < e; ++j) {
                const double dx = xi[0] - s.x[3 * j], dy = xi[1] - s.x[3 * j + 1],
                             dz = xi[2] - s.x[3 * j + 2];
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (r < p.h) {
                    const double a = h2 - r * r;
                    sum += k.c6 * a * a * a;
                }
            }
        });
        rho[i] = p.mass * sum;
    }
}

// Pairwise acceleration term of computeAccelerationCell (Solver.cu:231-309):
// pressure (spiky gradient), viscosity (Laplacian) and tension (Poly6), for
// 0 < r < h.  Accumulates the force fi that is later divided by rho_i.
inline void pair_force(const oracle_params& p, const Coeffs& k, const double* xi,
                       const double* vi, double rhoi, double pi,
                       const double* xj, const double* vj, double rhoj,
                       double* fi) {
    const double xij[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
    const double r = std::sqrt(xij[0] * xij[0] + xij[1] * xij[1] + xij[2] * xij[2]);
    if (!(r != 0.0 && r < p.h)) return;
    const double pj = p.stiffness * (rhoj - p.rest_density);
    const double coeffP = -rhoi * p.mass * (pi / (rhoi * rhoi) + pj / (rhoj * rhoj));
    const double gradA = -k.c45 * (p.h - r) * (p.h - r);   // spiky gradient magnitude
    const double coeffV = p.viscosity * p.mass / rhoj;
    const double lapl = k.c45 * (p.h - r);
    const double a = p.h * p.h - r * r;
    const double coeffT = -(k.c6 * a * a * a) * p.mass * p.tension;
    for (int d = 0; d < 3; ++d) {
        fi[d] += coeffP * gradA * xij[d] / r;
        fi[d] += coeffV * (vj[d] - vi[d]) * lapl;
        fi[d] += coeffT * xij[d];
    }
}

// computeAccelerationsD (Solver.cu:601-711) without boundary particles:
// a_i = f_i / rho_i + g.
void accelerations(const oracle_params& p, const State& s, const Neighbourhood& nb,
                   const std::vector<double>& rho, std::vector<double>& acc) {
    const uint32_t n = static_cast<uint32_t>(s.id.size());
    const Coeffs k(p.h);
    acc.assign(3 * static_cast<size_t>(n), 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        const double* xi = &s.x[3 * i];
        const double* vi = &s.v[3 * i];
        const double rhoi = rho[i];
        const double pi = p.stiffness * (rhoi - p.rest_density);
        double fi[3] = {0.0, 0.0, 0.0};
        for_each_stencil_cell(p, xi, [&](uint32_t c) {
            const uint32_t b = nb.start[c], e = nb.end[c];
            if (b == kEmpty) return;
            for (uint32_t j = b; j < e; ++j)
                pair_force(p, k, xi, vi, rhoi, pi, &s.x[3 * j], &s.v[3 * j], rho[j], fi);
        });
        for (int d = 0; d < 3; ++d) acc[3 * i + d] = fi[d] / rhoi + p.gravity[d];
    }
}

// integrateD (Solver.cu:713-755): v += dt a; x += dt v; plus the wall clamp
// of the benchmark scenes: outside [lo, hi] on an axis, x is put on the wall
// and that velocity component is reflected with restitution e.
void integrate(const oracle_params& p, State& s, const std::vector<double>& acc,
               double dt) {
    const size_t n = s.id.size();
    for (size_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            double v = s.v[3 * i + d] + dt * acc[3 * i + d];
            double x = s.x[3 * i + d] + dt * v;
            if (p.wall_clamp) {
                if (x < p.lo[d]) { x = p.lo[d]; v = -p.restitution * v; }
                if (x > p.hi[d]) { x = p.hi[d]; v = -p.restitution * v; }
            }
            s.v[3 * i + d] = v;
            s.x[3 * i + d] = x;
        }
    }
}

// evaluateBoundaryWeight (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:169-196).
double boundary_weight(double r, double h, double c) {
    const double q = 2.0 * r / h;
    const double coeff = 0.02 * c * c / r;
    if (q < 2.0 / 3.0) return coeff * 2.0 / 3.0;
    if (q < 1.0) return coeff * (2.0 * q - 1.5 * q * q);
    if (q < 2.0) return coeff * 0.5 * (2.0 - q) * (2.0 - q);
    return 0.0;
}

// Boundary particles sorted once on a grid, as the reference's constructor
// does (Solver.cu:1003-1059: hash, stable sort by hash, cell ranges).
struct WallSet {
    std::vector<double> x;
    Neighbourhood nb;
    double mass = 0.0, speed_sound = 0.0;
    uint32_t n = 0;
};

WallSet make_walls(const oracle_params& p, const oracle_walls& w) {
    WallSet W;
    W.n = w.n;
    W.mass = w.mass;
    W.speed_sound = w.speed_sound;
    State s;
    s.x.resize(3 * static_cast<size_t>(w.n));
    s.v.assign(3 * static_cast<size_t>(w.n), 0.0);
    s.id.resize(w.n);
    for (size_t i = 0; i < 3 * static_cast<size_t>(w.n); ++i) s.x[i] = w.pos[i];
    std::iota(s.id.begin(), s.id.end(), 0u);
    build_neighbourhood(p, s, W.nb);
    W.x = std::move(s.x);
    return W;
}

// computeBoundaryForceCell over the stencil of each particle
// (Solver.cu:311-344, 687-706): acc_i += B(r) m_b / (m + m_b) x_ij / r for
// 0 < r < h.  p gives the grid, h and the fluid mass m.
void add_wall_accel(const oracle_params& p, const WallSet& W, const std::vector<double>& x,
                    std::vector<double>& acc) {
    if (W.n == 0) return;
    const size_t n = x.size() / 3;
    const double ratio = W.mass / (p.mass + W.mass);
    for (size_t i = 0; i < n; ++i) {
        const double* xi = &x[3 * i];
        double b[3] = {0.0, 0.0, 0.0};
        for_each_stencil_cell(p, xi, [&](uint32_t c) {
            const uint32_t s = W.nb.start[c], e = W.nb.end[c];
            if (s == kEmpty) return;
            for (uint32_t j = s; j < e; ++j) {
                const double d[3] = {xi[0] - W.x[3 * j], xi[1] - W.x[3 * j + 1],
                                     xi[2] - W.x[3 * j + 2]};
                const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (!(r < p.h && r > 0.0)) continue;
                const double wgt = boundary_weight(r, p.h, W.speed_sound) * ratio;
                for (int k = 0; k < 3; ++k) b[k] += wgt * d[k] / r;
            }
        });
        for (int k = 0; k < 3; ++k) acc[3 * i + k] += b[k];
    }
}

// integrateXSPHD (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:346-395, 757-853) with
// each axis taking its own correction (the reference adds the x component to
// all three, :837-839): v' = v + dt a for every particle, then
// v'' = v' + eps sum_{r<h} (m / rho_j) W(r) (v'_j - v'_i), x' = x + dt v'',
// and the wall clamp.
void integrate_xsph(const oracle_params& p, State& s, const Neighbourhood& nb,
                    const std::vector<double>& rho, const std::vector<double>& acc, double dt) {
    const size_t n = s.id.size();
    const Coeffs k(p.h);
    std::vector<double> vp(3 * n);
    for (size_t i = 0; i < 3 * n; ++i) vp[i] = s.v[i] + dt * acc[i];
    std::vector<double> xo(3 * n), vo(3 * n);
    for (size_t i = 0; i < n; ++i) {
        const double* xi = &s.x[3 * i];
        double corr[3] = {0.0, 0.0, 0.0};
        for_each_stencil_cell(p, xi, [&](uint32_t c) {
            const uint32_t b = nb.start[c], e = nb.end[c];
            if (b == kEmpty) return;
            for (uint32_t j = b; j < e; ++j) {
                const double dx = xi[0] - s.x[3 * j], dy = xi[1] - s.x[3 * j + 1],
                             dz = xi[2] - s.x[3 * j + 2];
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (!(r < p.h)) continue;
                const double a = p.h * p.h - r * r;
                const double w = k.c6 * a * a * a * p.mass / rho[j];
                for (int d = 0; d < 3; ++d) corr[d] += (vp[3 * j + d] - vp[3 * i + d]) * w;
            }
        });
        for (int d = 0; d < 3; ++d) {
            double v = vp[3 * i + d] + p.xsph * corr[d];
            double x = xi[d] + dt * v;
            if (p.wall_clamp) {
                if (x < p.lo[d]) { x = p.lo[d]; v = -p.restitution * v; }
                if (x > p.hi[d]) { x = p.hi[d]; v = -p.restitution * v; }
            }
            vo[3 * i + d] = v;
            xo[3 * i + d] = x;
        }
    }
    s.x = std::move(xo);
    s.v = std::move(vo);
}

// evaluateMexicanHat3D (SPH2S_FINAL_TWO_SCALE/Solver.cu:206-213).
double mexican_hat(double x, double y, double z) {
    const double s = x * x + y * y + z * z;
    return 0.8673250705840776 * (s - 3.0) * std::exp(-0.5 * s);
}

State make_state(const float* pos, const float* vel, uint32_t n) {
    State s;
    s.x.resize(3 * static_cast<size_t>(n));
    s.v.assign(3 * static_cast<size_t>(n), 0.0);
    s.id.resize(n);
    for (size_t i = 0; i < 3 * static_cast<size_t>(n); ++i) {
        s.x[i] = pos[i];
        if (vel) s.v[i] = vel[i];
    }
    std::iota(s.id.begin(), s.id.end(), 0u);
    return s;
}

}  // namespace

extern "C" {

void oracle_hash(const oracle_grid* g, const float* pos, uint32_t n, uint32_t* out) {
    for (uint32_t i = 0; i < n; ++i)
        out[i] = hash_pos(*g, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
}

void oracle_stable_sort(const uint32_t* keys, uint32_t n, uint32_t* keys_out,
                        uint32_t* perm_out) {
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    for (uint32_t i = 0; i < n; ++i) {
        perm_out[i] = idx[i];
        keys_out[i] = keys[idx[i]];
    }
}

void oracle_cell_ranges(const uint32_t* k, uint32_t n, uint32_t C, uint32_t* start,
                        uint32_t* end) {
    for (uint32_t c = 0; c < C; ++c) start[c] = end[c] = kEmpty;
    for (uint32_t i = 0; i < n; ++i) {
        if (i == 0 || k[i] != k[i - 1]) {
            start[k[i]] = i;
            if (i > 0) end[k[i - 1]] = i;
        }
        if (i == n - 1) end[k[i]] = n;
    }
}

void oracle_energy_colors(const oracle_params* p, const float* pos, const float* vel, uint32_t n,
                          double* color_out, double* dist_out) {
    State s = make_state(pos, vel, n);
    Neighbourhood nb;
    build_neighbourhood(*p, s, nb);
    const double h = p->h, m = p->mass;
    for (uint32_t i = 0; i < n; ++i) {
        const double* xi = &s.x[3 * i];
        const double* vi = &s.v[3 * i];
        const double psi0 = mexican_hat(0.0, 0.0, 0.0);
        double psi_sum = psi0, msum = m;
        double w[3] = {vi[0] * psi0, vi[1] * psi0, vi[2] * psi0};
        double xc[3] = {m * xi[0], m * xi[1], m * xi[2]};
        for_each_stencil_cell(*p, xi, [&](uint32_t c) {
            const uint32_t b = nb.start[c], e = nb.end[c];
            if (b == kEmpty) return;
            for (uint32_t j = b; j < e; ++j) {
                const double d[3] = {xi[0] - s.x[3 * j], xi[1] - s.x[3 * j + 1],
                                     xi[2] - s.x[3 * j + 2]};
                const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (!(r != 0.0 && r < h)) continue;
                const double psi = mexican_hat(d[0] / h, d[1] / h, d[2] / h);
                for (int k = 0; k < 3; ++k) {
                    w[k] += s.v[3 * j + k] * psi;
                    xc[k] += m * s.x[3 * j + k];
                }
                psi_sum += psi;
                msum += m;
            }
        });
        double dd = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double t = xi[k] - xc[k] / msum;
            dd += t * t;
        }
        const double dist = std::sqrt(dd);
        const double ene = std::min((w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) /
                                        (psi_sum * psi_sum * h), 300.0);
        const uint32_t id = s.id[i];
        color_out[id] = (dist < 0.0025) ? 0.3 : ene / 300.0;
        if (dist_out) dist_out[id] = dist;
    }
}

void oracle_surface_indicator(const oracle_params* p_low, const float* pos_l, const double* rho_l,
                              uint32_t n_l, const float* pos_h, const double* rho_h, uint32_t n_h,
                              double* mag_out) {
    const double h0 = p_low->h, h1 = static_cast<double>(0.5f * p_low->h_f32);
    const double m0 = p_low->mass, m1 = static_cast<double>(static_cast<float>(p_low->mass) / 8.0f);
    const Coeffs k0(h0), k1(h1);
    auto spiky = [](double r, double h, double c45) { return c45 * (h - r) * (h - r) / r; };
    for (uint32_t i = 0; i < n_l; ++i) {
        double nv[3] = {0.0, 0.0, 0.0};
        auto add = [&](const float* xj, double wgt) {
            for (int k = 0; k < 3; ++k)
                nv[k] -= wgt * (static_cast<double>(pos_l[3 * i + k]) - xj[k]);
        };
        for (uint32_t j = 0; j < n_l; ++j) {
            double d2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double t = static_cast<double>(pos_l[3 * i + k]) - pos_l[3 * j + k];
                d2 += t * t;
            }
            const double r = std::sqrt(d2);
            if (r > 0.0 && r < h0) add(&pos_l[3 * j], m0 / rho_l[j] * spiky(r, h0, k0.c45));
        }
        for (uint32_t j = 0; j < n_h; ++j) {
            double d2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double t = static_cast<double>(pos_l[3 * i + k]) - pos_h[3 * j + k];
                d2 += t * t;
            }
            const double r = std::sqrt(d2);
            if (r > 0.0 && r < h0) {
                const double g = spiky(r, h0, k0.c45) + ((r < h1) ? spiky(r, h1, k1.c45) : 0.0);
                add(&pos_h[3 * j], m1 / rho_h[j] * 0.5 * g);
            }
        }
        mag_out[i] = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]) * h0;
    }
}

double oracle_boundary_weight(double r, double h, double speed_sound) {
    return boundary_weight(r, h, speed_sound);
}

void oracle_step(const oracle_params* p, const float* pos, const float* vel, uint32_t n,
                 double dt, uint32_t* hash_sorted, uint32_t* perm, uint32_t* cell_start,
                 uint32_t* cell_end, double* rho, double* acc, double* pos_out,
                 double* vel_out) {
    oracle_step_walls(p, nullptr, pos, vel, n, dt, hash_sorted, perm, cell_start, cell_end, rho,
                      acc, pos_out, vel_out);
}

void oracle_step_walls(const oracle_params* p, const oracle_walls* w, const float* pos,
                       const float* vel, uint32_t n, double dt, uint32_t* hash_sorted,
                       uint32_t* perm, uint32_t* cell_start, uint32_t* cell_end, double* rho,
                       double* acc, double* pos_out, double* vel_out) {
    State s = make_state(pos, vel, n);
    Neighbourhood nb;
    build_neighbourhood(*p, s, nb);
    std::vector<double> r, a;
    densities(*p, s, nb, r);
    accelerations(*p, s, nb, r, a);
    if (w && w->n) add_wall_accel(*p, make_walls(*p, *w), s.x, a);
    if (p->use_xsph)
        integrate_xsph(*p, s, nb, r, a, dt);
    else
        integrate(*p, s, a, dt);
    const uint32_t C = num_cells(p->grid);
    if (hash_sorted) std::copy(nb.hash_sorted.begin(), nb.hash_sorted.end(), hash_sorted);
    if (perm) std::copy(nb.perm.begin(), nb.perm.end(), perm);
    if (cell_start) std::copy(nb.start.begin(), nb.start.begin() + C, cell_start);
    if (cell_end) std::copy(nb.end.begin(), nb.end.begin() + C, cell_end);
    if (rho) std::copy(r.begin(), r.end(), rho);
    if (acc) std::copy(a.begin(), a.end(), acc);
    if (pos_out) std::copy(s.x.begin(), s.x.end(), pos_out);
    if (vel_out) std::copy(s.v.begin(), s.v.end(), vel_out);
}

void oracle_rollout(const oracle_params* p, const float* pos, const float* vel,
                    uint32_t n, double dt, uint32_t n_steps, double* pos_out,
                    double* vel_out, uint32_t* ids_out) {
    oracle_rollout_walls(p, nullptr, pos, vel, n, dt, n_steps, pos_out, vel_out, ids_out);
}

void oracle_rollout_walls(const oracle_params* p, const oracle_walls* w, const float* pos,
                          const float* vel, uint32_t n, double dt, uint32_t n_steps,
                          double* pos_out, double* vel_out, uint32_t* ids_out) {
    oracle_rollout_rho(p, w, pos, vel, n, dt, n_steps, pos_out, vel_out, ids_out, nullptr);
}

void oracle_rollout_rho(const oracle_params* p, const oracle_walls* w, const float* pos,
                        const float* vel, uint32_t n, double dt, uint32_t n_steps,
                        double* pos_out, double* vel_out, uint32_t* ids_out, double* rho_out) {
    State s = make_state(pos, vel, n);
    Neighbourhood nb;
    std::vector<double> r, a;
    WallSet W;
    if (w && w->n) W = make_walls(*p, *w);
    for (uint32_t t = 0; t < n_steps; ++t) {
        build_neighbourhood(*p, s, nb);
        densities(*p, s, nb, r);
        accelerations(*p, s, nb, r, a);
        add_wall_accel(*p, W, s.x, a);
        if (p->use_xsph)
            integrate_xsph(*p, s, nb, r, a, dt);
        else
            integrate(*p, s, a, dt);
    }
    if (pos_out) std::copy(s.x.begin(), s.x.end(), pos_out);
    if (vel_out) std::copy(s.v.begin(), s.v.end(), vel_out);
    if (ids_out) std::copy(s.id.begin(), s.id.end(), ids_out);
    if (rho_out) std::copy(r.begin(), r.end(), rho_out);
}

void oracle_density_bruteforce(const oracle_params* p, const double* x, uint32_t n,
                               double* rho) {
    const Coeffs k(p->h);
    for (uint32_t i = 0; i < n; ++i) {
        double sum = 0.0;
        for (uint32_t j = 0; j < n; ++j) {
            const double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1],
                         dz = x[3 * i + 2] - x[3 * j + 2];
            const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (r < p->h) {
                const double a = p->h * p->h - r * r;
                sum += k.c6 * a * a * a;
            }
        }
        rho[i] = p->mass * sum;
    }
}

void oracle_accel_bruteforce(const oracle_params* p, const double* x, const double* v,
                             const double* rho, uint32_t n, double* acc) {
    const Coeffs k(p->h);
    for (uint32_t i = 0; i < n; ++i) {
        const double rhoi = rho[i];
        const double pi = p->stiffness * (rhoi - p->rest_density);
        double fi[3] = {0.0, 0.0, 0.0};
        for (uint32_t j = 0; j < n; ++j)
            pair_force(*p, k, &x[3 * i], &v[3 * i], rhoi, pi, &x[3 * j], &v[3 * j], rho[j], fi);
        for (int d = 0; d < 3; ++d) acc[3 * i + d] = fi[d] / rhoi + p->gravity[d];
    }
}

double oracle_poly6(double r, double h) {
    if (!(r < h)) return 0.0;
    const Coeffs k(h);
    const double a = h * h - r * r;
    return k.c6 * a * a * a;
}

double oracle_spiky_grad_mag(double r, double h) {
    if (r == 0.0 || r > h) return 0.0;
    const Coeffs k(h);
    return -k.c45 * (h - r) * (h - r);
}

double oracle_visc_laplacian(double r, double h) {
    if (!(r < h)) return 0.0;
    const Coeffs k(h);
    return k.c45 * (h - r);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// FP32 OpenMP CPU step (the CPU baseline).  Same algorithm as the GPU path:
// FP32 hash, a stable counting sort by cell (which yields the cell ranges),
// reorder, density, force fused with symplectic Euler and the wall clamp.
// ---------------------------------------------------------------------------
struct cpu_solver {
    oracle_params p;
    uint32_t n = 0, C = 0;
    std::vector<float> x, v, tx, tv, rho;
    std::vector<uint32_t> id, tid, hash, count, start, end, perm;
};

extern "C" {

cpu_solver* cpu_solver_create(const oracle_params* p, const float* pos, const float* vel,
                              uint32_t n) {
    auto* s = new cpu_solver;
    s->p = *p;
    s->n = n;
    s->C = num_cells(p->grid);
    s->x.assign(pos, pos + 3 * static_cast<size_t>(n));
    s->v.assign(3 * static_cast<size_t>(n), 0.0f);
    if (vel) s->v.assign(vel, vel + 3 * static_cast<size_t>(n));
    s->id.resize(n);
    std::iota(s->id.begin(), s->id.end(), 0u);
    s->tx.resize(3 * static_cast<size_t>(n));
    s->tv.resize(3 * static_cast<size_t>(n));
    s->tid.resize(n);
    s->rho.resize(n);
    s->hash.resize(n);
    s->perm.resize(n);
    s->count.resize(s->C + 1);
    s->start.resize(s->C);
    s->end.resize(s->C);
    return s;
}

void cpu_solver_step(cpu_solver* s, float dt, uint32_t n_steps) {
    const oracle_params& p = s->p;
    const uint32_t n = s->n, C = s->C;
    const float h = p.h_f32, h2 = h * h;
    const Coeffs kd(static_cast<double>(h));
    const float c6m = static_cast<float>(kd.c6 * p.mass);
    const float c45 = static_cast<float>(kd.c45);
    const float m = static_cast<float>(p.mass), k = static_cast<float>(p.stiffness),
                rho0 = static_cast<float>(p.rest_density), mu = static_cast<float>(p.viscosity),
                sig = static_cast<float>(p.tension), e = static_cast<float>(p.restitution);
    const float c6 = static_cast<float>(kd.c6);
    const float g[3] = {static_cast<float>(p.gravity[0]), static_cast<float>(p.gravity[1]),
                        static_cast<float>(p.gravity[2])};
    const float lo[3] = {static_cast<float>(p.lo[0]), static_cast<float>(p.lo[1]),
                         static_cast<float>(p.lo[2])};
    const float hi[3] = {static_cast<float>(p.hi[0]), static_cast<float>(p.hi[1]),
                         static_cast<float>(p.hi[2])};
    for (uint32_t step = 0; step < n_steps; ++step) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < static_cast<int64_t>(n); ++i)
            s->hash[i] = hash_pos(p.grid, s->x[3 * i], s->x[3 * i + 1], s->x[3 * i + 2]);
        // stable counting sort by cell
        std::fill(s->count.begin(), s->count.end(), 0u);
        for (uint32_t i = 0; i < n; ++i) ++s->count[s->hash[i] + 1];
        for (uint32_t c = 0; c < C; ++c) s->count[c + 1] += s->count[c];
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < static_cast<int64_t>(C); ++c) {
            const bool empty = s->count[c] == s->count[c + 1];
            s->start[c] = empty ? kEmpty : s->count[c];
            s->end[c] = empty ? kEmpty : s->count[c + 1];
        }
        for (uint32_t i = 0; i < n; ++i) s->perm[s->count[s->hash[i]]++] = i;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
            const uint32_t j = s->perm[i];
            for (int d = 0; d < 3; ++d) {
                s->tx[3 * i + d] = s->x[3 * j + d];
                s->tv[3 * i + d] = s->v[3 * j + d];
            }
            s->tid[i] = s->id[j];
        }
        auto stencil = [&](const float* xi, auto&& f) {
            const oracle_grid& gr = p.grid;
            const int csx = coord_off(xi[0], -h, gr.origin[0], gr.spacing, gr.dims[0]);
            const int csy = coord_off(xi[1], -h, gr.origin[1], gr.spacing, gr.dims[1]);
            const int csz = coord_off(xi[2], -h, gr.origin[2], gr.spacing, gr.dims[2]);
            const int cex = coord_off(xi[0], h, gr.origin[0], gr.spacing, gr.dims[0]);
            const int cey = coord_off(xi[1], h, gr.origin[1], gr.spacing, gr.dims[1]);
            const int cez = coord_off(xi[2], h, gr.origin[2], gr.spacing, gr.dims[2]);
            for (int cz = csz; cz <= cez; ++cz)
                for (int cy = csy; cy <= cey; ++cy) {
                    uint32_t b = kEmpty, en = 0;
                    for (int cx = csx; cx <= cex; ++cx) {
                        const uint32_t c = hash_of(gr, cx, cy, cz);
                        if (s->start[c] != kEmpty) {
                            b = std::min(b, s->start[c]);
                            en = std::max(en, s->end[c]);
                        }
                    }
                    for (uint32_t j = b; j < en; ++j) f(j);
                }
        };
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
            const float* xi = &s->tx[3 * i];
            float sum = 0.0f;
            stencil(xi, [&](uint32_t j) {
                const float dx = xi[0] - s->tx[3 * j], dy = xi[1] - s->tx[3 * j + 1],
                            dz = xi[2] - s->tx[3 * j + 2];
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 < h2) {
                    const float a = h2 - r2;
                    sum += a * a * a;
                }
            });
            s->rho[i] = c6m * sum;
        }
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
            const float* xi = &s->tx[3 * i];
            const float* vi = &s->tv[3 * i];
            const float rhoi = s->rho[i];
            const float Ai = k * (rhoi - rho0) / (rhoi * rhoi);
            float P[3] = {0, 0, 0}, V[3] = {0, 0, 0}, T[3] = {0, 0, 0};
            stencil(xi, [&](uint32_t j) {
                const float d[3] = {xi[0] - s->tx[3 * j], xi[1] - s->tx[3 * j + 1],
                                    xi[2] - s->tx[3 * j + 2]};
                const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (r2 < h2 && r2 > 0.0f) {
                    const float r = std::sqrt(r2);
                    const float hr = h - r;
                    const float rhoj = s->rho[j];
                    const float Aj = k * (rhoj - rho0) / (rhoj * rhoj);
                    const float tp = (Ai + Aj) * hr * hr / r;
                    const float tv = hr / rhoj;
                    const float a = h2 - r2;
                    const float a3 = a * a * a;
                    for (int q = 0; q < 3; ++q) {
                        P[q] += tp * d[q];
                        V[q] += tv * (s->tv[3 * j + q] - vi[q]);
                        T[q] += a3 * d[q];
                    }
                }
            });
            for (int q = 0; q < 3; ++q) {
                const float acc = m * c45 * P[q] + mu * m * c45 / rhoi * V[q] -
                                  sig * m * c6 / rhoi * T[q] + g[q];
                float vq = vi[q] + dt * acc;
                float xq = xi[q] + dt * vq;
                if (p.wall_clamp) {
                    if (xq < lo[q]) { xq = lo[q]; vq = -e * vq; }
                    if (xq > hi[q]) { xq = hi[q]; vq = -e * vq; }
                }
                s->x[3 * i + q] = xq;
                s->v[3 * i + q] = vq;
            }
        }
        std::swap(s->id, s->tid);
    }
}

void cpu_solver_get(const cpu_solver* s, float* pos_out, float* vel_out, float* rho_out,
                    uint32_t* ids_out) {
    if (pos_out) std::copy(s->x.begin(), s->x.end(), pos_out);
    if (vel_out) std::copy(s->v.begin(), s->v.end(), vel_out);
    if (rho_out) std::copy(s->rho.begin(), s->rho.end(), rho_out);
    if (ids_out) std::copy(s->id.begin(), s->id.end(), ids_out);
}

int cpu_solver_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void cpu_solver_destroy(cpu_solver* s) { delete s; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Two-scale step (FP64).  Restates TWO_SCALE/Solver.cu:1568-1614 under two
// couplings.  Fixed (coupling 0, the HIP default): complementary
// averaged-kernel sums (TWO_SCALE/Solver.cu:252-290, 394-499) with blend
// weight 1, split of low particles inside the region into 8 children, merge of
// sibling groups that are all outside the region expanded by h0/2.  Blended
// (coupling 1): every neighbour term of both the own-set and the complementary
// sums is weighted by the neighbour's blend weight lambda_j
// (TWO_SCALE/Solver.cu:215-499), and integrateD's state machine runs after
// the integration (Solver.cu:1240-1289): state & 3 is the mode (0 settled,
// 1 fading out, 2 fading in, 3 split requested), state & 4 marks the high set;
// lambda moves by -inc / +inc in modes 1 / 2 (FP32 arithmetic, as the
// reference keeps lambda in float); a particle in mode 1 with lambda <= 0 is
// removed; one in mode 2 reaching lambda >= 1 settles (mode 0, lambda = 1).
// A low particle in mode 3, or settled and inside the region, splits: 8
// children with lambda 0 in mode 2 (state 6), and the parent fades out
// (state 1) instead of being removed.  The faded merge is the reverse: a group
// of 8 settled siblings outside the expanded region emits a low particle at
// their centroid with lambda 0 in mode 2, and the siblings fade out.  A split
// waits while a live child of the same group exists, and a merge while a live
// low particle with the group's id exists, so ids stay unique.
// ---------------------------------------------------------------------------
namespace {

struct TsSet {
    oracle_params p;
    std::vector<double> x, v;            // state, AoS
    std::vector<uint32_t> id, tag;
    std::vector<float> lam;              // blend weight lambda
    std::vector<uint32_t> state;         // state word
    // last step, sorted order
    std::vector<double> rho, acc;
    Neighbourhood nb;
    uint32_t n_step = 0;
};

template <class F>
void for_each_cell_in(const oracle_grid& g, float hs, const double* xi, F&& f) {
    const float x = static_cast<float>(xi[0]), y = static_cast<float>(xi[1]),
                z = static_cast<float>(xi[2]);
    const int csx = coord_off(x, -hs, g.origin[0], g.spacing, g.dims[0]);
    const int csy = coord_off(y, -hs, g.origin[1], g.spacing, g.dims[1]);
    const int csz = coord_off(z, -hs, g.origin[2], g.spacing, g.dims[2]);
    const int cex = coord_off(x, hs, g.origin[0], g.spacing, g.dims[0]);
    const int cey = coord_off(y, hs, g.origin[1], g.spacing, g.dims[1]);
    const int cez = coord_off(z, hs, g.origin[2], g.spacing, g.dims[2]);
    for (int cz = csz; cz <= cez; ++cz)
        for (int cy = csy; cy <= cey; ++cy)
            for (int cx = csx; cx <= cex; ++cx) f(hash_of(g, cx, cy, cz));
}

void sort_set(TsSet& S) {
    State st;
    st.x = S.x;
    st.v = S.v;
    st.id = S.id;
    const uint32_t n = static_cast<uint32_t>(S.id.size());
    build_neighbourhood(S.p, st, S.nb);
    std::vector<uint32_t> tag(n), state(n);
    std::vector<float> lam(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t j = S.nb.perm[i];
        tag[i] = S.tag[j];
        lam[i] = S.lam[j];
        state[i] = S.state[j];
    }
    S.x = std::move(st.x);
    S.v = std::move(st.v);
    S.id = std::move(st.id);
    S.tag = std::move(tag);
    S.lam = std::move(lam);
    S.state = std::move(state);
    S.n_step = n;
}

// v += dt a, the velocity cap |v| <= vmax (vmax > 0; Solver.cu:1212-1223,
// without the reference's 0/0 at |v| = 0), x += dt v, then the wall clamp.
void integrate_ts(const oracle_params& box, TsSet& A, double dt, float vmax) {
    const size_t n = A.id.size();
    for (size_t i = 0; i < n; ++i) {
        double v[3], x[3];
        for (int d = 0; d < 3; ++d) v[d] = A.v[3 * i + d] + dt * A.acc[3 * i + d];
        if (vmax > 0.0f) {
            const double nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            if (nv > vmax)
                for (int d = 0; d < 3; ++d) v[d] *= static_cast<double>(vmax) / nv;
        }
        for (int d = 0; d < 3; ++d) {
            x[d] = A.x[3 * i + d] + dt * v[d];
            if (box.wall_clamp) {
                if (x[d] < box.lo[d]) { x[d] = box.lo[d]; v[d] = -box.restitution * v[d]; }
                if (x[d] > box.hi[d]) { x[d] = box.hi[d]; v[d] = -box.restitution * v[d]; }
            }
            A.v[3 * i + d] = v[d];
            A.x[3 * i + d] = x[d];
        }
    }
}

}  // namespace

struct oracle_ts {
    TsSet set[2];
    float rlo[3], rhi[3], margin_f, child_d_f;
    uint32_t max_low = 0;
    oracle_ts_options opt{0u, 0u, 1u, 0.0125f, 0.0f};
    std::vector<double> split_rho;   // low densities the split offsets use (sorted order)
    WallSet walls;                   // boundary particles on the low grid
};

namespace {

inline bool ts_blended(const oracle_ts* t) { return t->opt.coupling == 1u; }

// The mistake planted by oracle_ts_plant_mistake (0 = none); see sph_oracle.h.
int g_planted = 0;

// Fixed coupling: own-set density and acceleration of the single-scale step,
// plus the complementary sums with blend weight 1.
void ts_forces_fixed(oracle_ts* t) {
    TsSet& L = t->set[0];
    TsSet& H = t->set[1];
    const double h0 = L.p.h, h1 = H.p.h;
    const Coeffs k0(h0), k1(h1);
    for (int r = 0; r < 2; ++r) {
        TsSet& S = t->set[r];
        State st;
        st.x = S.x;
        st.v = S.v;
        st.id = S.id;
        densities(S.p, st, S.nb, S.rho);
    }
    // complementary densities: + m_B sum_{r < h0} 1/2 (W_h0 + W_h1)
    const double half = (g_planted == 3) ? 1.0 : 0.5;
    const double cut = (g_planted == 2) ? h1 : h0;
    const float reach = (g_planted == 4) ? H.p.h_f32 : L.p.h_f32;
    std::vector<double> rho_new[2];
    for (int r = 0; r < 2; ++r) {
        TsSet& A = t->set[r];
        TsSet& B = t->set[1 - r];
        const uint32_t na = static_cast<uint32_t>(A.id.size());
        rho_new[r] = A.rho;
        for (uint32_t i = 0; i < na; ++i) {
            const double* xi = &A.x[3 * i];
            double sum = 0.0;
            if (!B.id.empty())
                for_each_cell_in(B.p.grid, reach, xi, [&](uint32_t c) {
                    const uint32_t b = B.nb.start[c], e = B.nb.end[c];
                    if (b == kEmpty) return;
                    for (uint32_t j = b; j < e; ++j) {
                        const double dx = xi[0] - B.x[3 * j], dy = xi[1] - B.x[3 * j + 1],
                                     dz = xi[2] - B.x[3 * j + 2];
                        const double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
                        if (rr < cut) {
                            const double a0 = h0 * h0 - rr * rr, a1 = h1 * h1 - rr * rr;
                            const double w0 = k0.c6 * a0 * a0 * a0;
                            const double w1 = (rr < h1) ? k1.c6 * a1 * a1 * a1 : 0.0;
                            sum += half * (w0 + w1);
                        }
                    }
                });
            rho_new[r][i] += ((g_planted == 1) ? A.p.mass : B.p.mass) * sum;
        }
    }
    L.rho = rho_new[0];
    H.rho = rho_new[1];
    for (int r = 0; r < 2; ++r) {
        TsSet& A = t->set[r];
        TsSet& B = t->set[1 - r];
        State st;
        st.x = A.x;
        st.v = A.v;
        st.id = A.id;
        accelerations(A.p, st, A.nb, A.rho, A.acc);
        const uint32_t na = static_cast<uint32_t>(A.id.size());
        for (uint32_t i = 0; i < na; ++i) {
            const double* xi = &A.x[3 * i];
            const double* vi = &A.v[3 * i];
            const double rhoi = A.rho[i];
            const double Ai = A.p.stiffness * (rhoi - A.p.rest_density) / (rhoi * rhoi);
            double f[3] = {0.0, 0.0, 0.0};
            const double mB = (g_planted == 1) ? A.p.mass : B.p.mass;
            if (!B.id.empty())
                for_each_cell_in(B.p.grid, reach, xi, [&](uint32_t c) {
                    const uint32_t b = B.nb.start[c], e = B.nb.end[c];
                    if (b == kEmpty) return;
                    for (uint32_t j = b; j < e; ++j) {
                        const double xij[3] = {xi[0] - B.x[3 * j], xi[1] - B.x[3 * j + 1],
                                               xi[2] - B.x[3 * j + 2]};
                        const double rr = std::sqrt(xij[0] * xij[0] + xij[1] * xij[1] +
                                                    xij[2] * xij[2]);
                        if (!(rr != 0.0 && rr < cut)) continue;
                        const double rhoj = B.rho[j];
                        const double Aj = A.p.stiffness * (rhoj - A.p.rest_density) / (rhoj * rhoj);
                        const double hr0 = h0 - rr, hr1 = std::max(h1 - rr, 0.0);
                        const double g = half * (k0.c45 * hr0 * hr0 + k1.c45 * hr1 * hr1) / rr;
                        const double l = half * (k0.c45 * hr0 + k1.c45 * hr1);
                        const double a0 = h0 * h0 - rr * rr, a1 = h1 * h1 - rr * rr;
                        const double w = half * (k0.c6 * a0 * a0 * a0 +
                                                 ((rr < h1) ? k1.c6 * a1 * a1 * a1 : 0.0));
                        for (int d = 0; d < 3; ++d) {
                            f[d] += mB * (Ai + Aj) * g * xij[d];
                            f[d] += A.p.viscosity * mB / (rhoi * rhoj) * l *
                                    (B.v[3 * j + d] - vi[d]);
                            f[d] -= A.p.tension * mB / rhoi * w * xij[d];
                        }
                    }
                });
            for (int d = 0; d < 3; ++d) A.acc[3 * i + d] += f[d];
        }
    }
}

// Blended coupling: computeDensitiesPressuresD and computeAccelerationsD of
// TWO_SCALE/Solver.cu:732-1048 (without the boundary particles, whose role
// the wall clamp takes): every term weighted by the neighbour's lambda.
void ts_forces_blended(oracle_ts* t) {
    TsSet& L = t->set[0];
    TsSet& H = t->set[1];
    const double h0 = L.p.h, h1 = H.p.h;
    const Coeffs k0(h0), k1(h1);
    // densities: m_A sum_A lambda_j W_hA + m_B sum_{B, r < h0} lambda_j 1/2 (W_h0 + W_h1)
    for (int r = 0; r < 2; ++r) {
        TsSet& A = t->set[r];
        TsSet& B = t->set[1 - r];
        const Coeffs ka(A.p.h);
        const double ha = A.p.h;
        const uint32_t na = static_cast<uint32_t>(A.id.size());
        A.rho.assign(na, 0.0);
        for (uint32_t i = 0; i < na; ++i) {
            const double* xi = &A.x[3 * i];
            double own = 0.0, cross = 0.0;
            for_each_stencil_cell(A.p, xi, [&](uint32_t c) {
                const uint32_t b = A.nb.start[c], e = A.nb.end[c];
                if (b == kEmpty) return;
                for (uint32_t j = b; j < e; ++j) {
                    const double dx = xi[0] - A.x[3 * j], dy = xi[1] - A.x[3 * j + 1],
                                 dz = xi[2] - A.x[3 * j + 2];
                    const double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
                    if (rr < ha) {
                        const double a = ha * ha - rr * rr;
                        own += static_cast<double>(A.lam[j]) * ka.c6 * a * a * a;
                    }
                }
            });
            if (!B.id.empty())
                for_each_cell_in(B.p.grid, L.p.h_f32, xi, [&](uint32_t c) {
                    const uint32_t b = B.nb.start[c], e = B.nb.end[c];
                    if (b == kEmpty) return;
                    for (uint32_t j = b; j < e; ++j) {
                        const double dx = xi[0] - B.x[3 * j], dy = xi[1] - B.x[3 * j + 1],
                                     dz = xi[2] - B.x[3 * j + 2];
                        const double rr = std::sqrt(dx * dx + dy * dy + dz * dz);
                        if (rr < h0) {
                            const double a0 = h0 * h0 - rr * rr, a1 = h1 * h1 - rr * rr;
                            const double w0 = k0.c6 * a0 * a0 * a0;
                            const double w1 = (rr < h1) ? k1.c6 * a1 * a1 * a1 : 0.0;
                            cross += static_cast<double>(B.lam[j]) * 0.5 * (w0 + w1);
                        }
                    }
                });
            A.rho[i] = A.p.mass * own + B.p.mass * cross;
        }
    }
    // accelerations: f_i / rho_i + g, f of computeAccelerationCell(Compl)
    for (int r = 0; r < 2; ++r) {
        TsSet& A = t->set[r];
        TsSet& B = t->set[1 - r];
        const Coeffs ka(A.p.h);
        const uint32_t na = static_cast<uint32_t>(A.id.size());
        A.acc.assign(3 * static_cast<size_t>(na), 0.0);
        for (uint32_t i = 0; i < na; ++i) {
            const double* xi = &A.x[3 * i];
            const double* vi = &A.v[3 * i];
            const double rhoi = A.rho[i];
            const double pi = A.p.stiffness * (rhoi - A.p.rest_density);
            double f[3] = {0.0, 0.0, 0.0};
            for_each_stencil_cell(A.p, xi, [&](uint32_t c) {
                const uint32_t b = A.nb.start[c], e = A.nb.end[c];
                if (b == kEmpty) return;
                for (uint32_t j = b; j < e; ++j) {
                    double fj[3] = {0.0, 0.0, 0.0};
                    pair_force(A.p, ka, xi, vi, rhoi, pi, &A.x[3 * j], &A.v[3 * j], A.rho[j], fj);
                    for (int d = 0; d < 3; ++d) f[d] += static_cast<double>(A.lam[j]) * fj[d];
                }
            });
            if (!B.id.empty())
                for_each_cell_in(B.p.grid, L.p.h_f32, xi, [&](uint32_t c) {
                    const uint32_t b = B.nb.start[c], e = B.nb.end[c];
                    if (b == kEmpty) return;
                    for (uint32_t j = b; j < e; ++j) {
                        const double xij[3] = {xi[0] - B.x[3 * j], xi[1] - B.x[3 * j + 1],
                                               xi[2] - B.x[3 * j + 2]};
                        const double rr = std::sqrt(xij[0] * xij[0] + xij[1] * xij[1] +
                                                    xij[2] * xij[2]);
                        if (!(rr != 0.0 && rr < h0)) continue;
                        const double lj = static_cast<double>(B.lam[j]);
                        const double rhoj = B.rho[j];
                        const double pj = A.p.stiffness * (rhoj - A.p.rest_density);
                        // Solver.cu:394-499: averaged spiky gradient, Laplacian and Poly6.
                        const double coeffP =
                            -rhoi * B.p.mass * (pi / (rhoi * rhoi) + pj / (rhoj * rhoj));
                        const double gmag = -0.5 * (k0.c45 * (h0 - rr) * (h0 - rr) +
                                                    ((rr < h1) ? k1.c45 * (h1 - rr) * (h1 - rr)
                                                               : 0.0));
                        const double coeffV = A.p.viscosity * B.p.mass / rhoj;
                        const double lap = 0.5 * (k0.c45 * (h0 - rr) +
                                                  ((rr < h1) ? k1.c45 * (h1 - rr) : 0.0));
                        const double a0 = h0 * h0 - rr * rr, a1 = h1 * h1 - rr * rr;
                        const double w = 0.5 * (k0.c6 * a0 * a0 * a0 +
                                                ((rr < h1) ? k1.c6 * a1 * a1 * a1 : 0.0));
                        const double coeffT = -w * B.p.mass * A.p.tension;
                        for (int d = 0; d < 3; ++d) {
                            f[d] += lj * coeffP * gmag * xij[d] / rr;
                            f[d] += lj * coeffV * (B.v[3 * j + d] - vi[d]) * lap;
                            f[d] += lj * coeffT * xij[d];
                        }
                    }
                });
            for (int d = 0; d < 3; ++d) A.acc[3 * i + d] = f[d] / rhoi + A.p.gravity[d];
        }
    }
}

// integrateD's blend update (Solver.cu:1240-1262) in FP32.
void blend_update(TsSet& A, float inc) {
    for (size_t i = 0; i < A.id.size(); ++i) {
        const uint32_t mode = A.state[i] & 3u;
        const float c = (mode == 1u) ? -1.0f : (mode == 2u) ? 1.0f : 0.0f;
        float lam = A.lam[i] + c * inc;
        if (mode == 2u && lam >= 1.0f) {
            A.state[i] &= 4u;
            lam = 1.0f;
        }
        A.lam[i] = lam;
    }
}

}  // namespace

extern "C" {

oracle_ts* oracle_ts_create(const oracle_params* p_low, const oracle_grid* grid_high,
                            const float region_lo[3], const float region_hi[3],
                            uint32_t max_low) {
    auto* t = new oracle_ts;
    t->set[0].p = *p_low;
    t->set[1].p = *p_low;
    oracle_params& ph = t->set[1].p;
    ph.h_f32 = 0.5f * p_low->h_f32;
    ph.h = static_cast<double>(ph.h_f32);
    ph.mass = static_cast<double>(static_cast<float>(p_low->mass) / 8.0f);
    ph.grid = *grid_high;
    for (int d = 0; d < 3; ++d) {
        t->rlo[d] = region_lo[d];
        t->rhi[d] = region_hi[d];
    }
    t->margin_f = 0.5f * p_low->h_f32;
    t->child_d_f = 0.25f * ph.h_f32;
    t->max_low = max_low;
    return t;
}

void oracle_ts_destroy(oracle_ts* t) { delete t; }

void oracle_ts_plant_mistake(int which) { g_planted = which; }

void oracle_ts_set_state(oracle_ts* t, const float* pos_l, const float* vel_l,
                         const uint32_t* ids_l, uint32_t n_l, const float* pos_h,
                         const float* vel_h, const uint32_t* ids_h, const uint32_t* tags_h,
                         uint32_t n_h) {
    const float* P[2] = {pos_l, pos_h};
    const float* V[2] = {vel_l, vel_h};
    const uint32_t* I[2] = {ids_l, ids_h};
    const uint32_t N[2] = {n_l, n_h};
    for (int r = 0; r < 2; ++r) {
        TsSet& S = t->set[r];
        const uint32_t n = N[r];
        S.x.assign(3 * static_cast<size_t>(n), 0.0);
        S.v.assign(3 * static_cast<size_t>(n), 0.0);
        S.id.resize(n);
        S.tag.assign(n, 0u);
        S.lam.assign(n, 1.0f);
        S.state.assign(n, r == 0 ? 0u : 4u);
        for (size_t i = 0; i < 3 * static_cast<size_t>(n); ++i) {
            S.x[i] = P[r][i];
            if (V[r]) S.v[i] = V[r][i];
        }
        for (uint32_t i = 0; i < n; ++i) {
            S.id[i] = I[r] ? I[r][i] : i;
            if (r == 1) S.tag[i] = tags_h[i];
        }
    }
}

void oracle_ts_set_options(oracle_ts* t, const oracle_ts_options* o) {
    const bool was = ts_blended(t);
    t->opt = *o;
    if (ts_blended(t) && !was)
        for (int r = 0; r < 2; ++r) {
            TsSet& S = t->set[r];
            S.lam.assign(S.id.size(), 1.0f);
            S.state.assign(S.id.size(), r == 0 ? 0u : 4u);
        }
}

void oracle_ts_set_blend(oracle_ts* t, int res, const float* lambda, const uint32_t* state) {
    TsSet& S = t->set[res];
    for (size_t i = 0; i < S.id.size(); ++i) {
        if (lambda) S.lam[i] = lambda[i];
        if (state) S.state[i] = state[i];
    }
}

void oracle_ts_set_walls(oracle_ts* t, const oracle_walls* w) {
    t->walls = (w && w->n) ? make_walls(t->set[0].p, *w) : WallSet{};
}

void oracle_ts_mark_split(oracle_ts* t, const uint32_t* ids, uint32_t n) {
    TsSet& L = t->set[0];
    std::vector<uint32_t> want(ids, ids + n);
    std::sort(want.begin(), want.end());
    for (size_t i = 0; i < L.id.size(); ++i)
        if ((L.state[i] & 7u) == 0u && std::binary_search(want.begin(), want.end(), L.id[i]))
            L.state[i] |= 3u;
}

void oracle_ts_step(oracle_ts* t, double dt) {
    TsSet& L = t->set[0];
    TsSet& H = t->set[1];
    sort_set(H);
    sort_set(L);
    if (ts_blended(t))
        ts_forces_blended(t);
    else
        ts_forces_fixed(t);
    // boundary particles: low grid, h0 and m0 for both sets
    for (int r = 0; r < 2; ++r) add_wall_accel(L.p, t->walls, t->set[r].x, t->set[r].acc);
    for (int r = 0; r < 2; ++r) {
        TsSet& A = t->set[r];
        // 0.8f * h / dt of Solver.cu:1214-1217, rounded as the FP32 step does.
        const float vmax = (t->opt.velocity_cap > 0.0f)
                               ? t->opt.velocity_cap * A.p.h_f32 / static_cast<float>(dt)
                               : 0.0f;
        integrate_ts(L.p, A, dt, vmax);  // clamp box and restitution of the low configuration
        if (ts_blended(t)) blend_update(A, t->opt.blend_increment);
    }
    t->split_rho = L.rho;
    oracle_ts_split_merge(t);
}

void oracle_ts_split_merge(oracle_ts* t) {
    TsSet& L = t->set[0];
    TsSet& H = t->set[1];
    const bool blended = ts_blended(t);
    const uint32_t nl = static_cast<uint32_t>(L.id.size()), nh = static_cast<uint32_t>(H.id.size());
    // The split offsets of a standalone pass (no step before it) use rest density.
    if (t->split_rho.size() != nl) t->split_rho.assign(nl, L.p.rest_density);
    const float m = t->margin_f;
    const float olo[3] = {t->rlo[0] - m, t->rlo[1] - m, t->rlo[2] - m};
    const float ohi[3] = {t->rhi[0] + m, t->rhi[1] + m, t->rhi[2] + m};
    auto inside = [&](const double* x) {
        for (int d = 0; d < 3; ++d)
            if (!(x[d] >= t->rlo[d] && x[d] <= t->rhi[d])) return false;
        return true;
    };
    auto outside_margin = [&](const double* x) {
        for (int d = 0; d < 3; ++d)
            if (x[d] < olo[d] || x[d] > ohi[d]) return true;
        return false;
    };
    auto removed = [&](const TsSet& S, uint32_t i) {
        return blended && (S.state[i] & 3u) == 1u && S.lam[i] <= 0.0f;
    };
    constexpr uint32_t kMask = (1u << 29) - 1u;
    // Per group: settled children outside the expanded region, and whether a
    // live child (bit 0) or a live low particle with the group's id (bit 1) exists.
    std::vector<uint32_t> cnt(t->max_low, 0u), live(t->max_low, 0u);
    std::vector<double> gx(static_cast<size_t>(t->max_low) * 24, 0.0),
        gv(static_cast<size_t>(t->max_low) * 24, 0.0);
    for (uint32_t i = 0; i < nh; ++i) {
        const uint32_t g = H.tag[i] & kMask, k = H.tag[i] >> 29;
        if (g >= t->max_low || removed(H, i)) continue;
        live[g] |= 1u;
        if (!t->opt.merge || !outside_margin(&H.x[3 * i])) continue;
        if (blended && (H.state[i] & 3u) != 0u) continue;
        ++cnt[g];
        for (int d = 0; d < 3; ++d) {
            gx[(static_cast<size_t>(g) * 8 + k) * 3 + d] = H.x[3 * i + d];
            gv[(static_cast<size_t>(g) * 8 + k) * 3 + d] = H.v[3 * i + d];
        }
    }
    TsSet nL, nH;
    auto push = [](TsSet& S, const double* x, const double* v, uint32_t id, uint32_t tag,
                   float lam, uint32_t state) {
        for (int d = 0; d < 3; ++d) {
            S.x.push_back(x[d]);
            S.v.push_back(v[d]);
        }
        S.id.push_back(id);
        S.tag.push_back(tag);
        S.lam.push_back(lam);
        S.state.push_back(state);
    };
    std::vector<uint32_t> split;
    for (uint32_t i = 0; i < nl; ++i) {
        if (removed(L, i)) continue;
        live[L.id[i]] |= 2u;
        const uint32_t mode = L.state[i] & 3u;
        const bool s = blended ? ((mode == 3u || (mode == 0u && inside(&L.x[3 * i]))) &&
                                  !(live[L.id[i]] & 1u))
                               : inside(&L.x[3 * i]);
        if (s) split.push_back(i);
        if (!s || blended)
            push(nL, &L.x[3 * i], &L.v[3 * i], L.id[i], 0u, L.lam[i], s ? 1u : L.state[i]);
    }
    std::vector<uint32_t> leaders;
    for (uint32_t i = 0; i < nh; ++i) {
        if (removed(H, i)) continue;
        const uint32_t g = H.tag[i] & kMask, k = H.tag[i] >> 29;
        const bool merged = g < t->max_low && cnt[g] == 8u && !(blended && (live[g] & 2u));
        if (!merged || blended)
            push(nH, &H.x[3 * i], &H.v[3 * i], H.id[i], H.tag[i], H.lam[i],
                 merged ? ((H.state[i] & 4u) | 1u) : H.state[i]);
        if (merged && k == 0u) leaders.push_back(g);
    }
    for (uint32_t g : leaders) {
        double x[3] = {0, 0, 0}, v[3] = {0, 0, 0};
        for (int k = 0; k < 8; ++k)
            for (int d = 0; d < 3; ++d) {
                x[d] += gx[(static_cast<size_t>(g) * 8 + k) * 3 + d];
                v[d] += gv[(static_cast<size_t>(g) * 8 + k) * 3 + d];
            }
        for (int d = 0; d < 3; ++d) {
            x[d] /= 8.0;
            v[d] /= 8.0;
        }
        push(nL, x, v, g, 0u, blended ? 0.0f : 1.0f, blended ? 2u : 0u);
    }
    const double cd = static_cast<double>(t->child_d_f);
    for (uint32_t i : split) {
        const uint32_t pid = L.id[i];
        // insertHighResParticles (Solver.cu:541-586): r/sqrt(3) along the
        // diagonals with r = 1/2 (3 m0 / (4 pi rho))^(1/3); else +/- h1/4.
        const double rr = 0.5 * std::cbrt(3.0 / (4.0 * kPi) * L.p.mass / t->split_rho[i]);
        const double d_ref = rr / std::sqrt(3.0);
        for (uint32_t k = 0; k < 8; ++k) {
            double x[3];
            for (int d = 0; d < 3; ++d) {
                const double off = t->opt.split_offsets ? d_ref : cd;
                x[d] = L.x[3 * i + d] + (((k >> d) & 1u) ? off : -off);
            }
            push(nH, x, &L.v[3 * i], t->max_low + 8u * pid + k, (k << 29) | (pid & kMask),
                 blended ? 0.0f : 1.0f, blended ? 6u : 4u);
        }
    }
    L.x = std::move(nL.x);
    L.v = std::move(nL.v);
    L.id = std::move(nL.id);
    L.tag = std::move(nL.tag);
    L.lam = std::move(nL.lam);
    L.state = std::move(nL.state);
    H.x = std::move(nH.x);
    H.v = std::move(nH.v);
    H.id = std::move(nH.id);
    H.tag = std::move(nH.tag);
    H.lam = std::move(nH.lam);
    H.state = std::move(nH.state);
    t->split_rho.clear();
}

void oracle_ts_counts(const oracle_ts* t, uint32_t counts[4]) {
    counts[0] = static_cast<uint32_t>(t->set[0].id.size());
    counts[1] = static_cast<uint32_t>(t->set[1].id.size());
    counts[2] = t->set[0].n_step;
    counts[3] = t->set[1].n_step;
}

void oracle_ts_get(const oracle_ts* t, int res, int field, void* out) {
    const TsSet& S = t->set[res];
    switch (field) {
        case 0: std::copy(S.x.begin(), S.x.end(), static_cast<double*>(out)); break;
        case 1: std::copy(S.v.begin(), S.v.end(), static_cast<double*>(out)); break;
        case 2: std::copy(S.id.begin(), S.id.end(), static_cast<uint32_t*>(out)); break;
        case 3: std::copy(S.tag.begin(), S.tag.end(), static_cast<uint32_t*>(out)); break;
        case 4: std::copy(S.rho.begin(), S.rho.end(), static_cast<double*>(out)); break;
        case 5: std::copy(S.acc.begin(), S.acc.end(), static_cast<double*>(out)); break;
        case 6: std::copy(S.nb.hash_sorted.begin(), S.nb.hash_sorted.end(), static_cast<uint32_t*>(out)); break;
        case 7: std::copy(S.nb.perm.begin(), S.nb.perm.end(), static_cast<uint32_t*>(out)); break;
        case 8: std::copy(S.nb.start.begin(), S.nb.start.end(), static_cast<uint32_t*>(out)); break;
        case 9: std::copy(S.nb.end.begin(), S.nb.end.end(), static_cast<uint32_t*>(out)); break;
        case 10: std::copy(S.lam.begin(), S.lam.end(), static_cast<float*>(out)); break;
        case 11: std::copy(S.state.begin(), S.state.end(), static_cast<uint32_t*>(out)); break;
        default: break;
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Differentiable rollout (FP64): reverse-mode adjoint of n steps of the
// single-scale step above with the pair topology frozen within each step.
// Per step, with A_i = k (rho_i - rho0) / rho_i^2, B_i = 1 / rho_i,
// F = (h - r)^2 / r, C45 = m c45, Vc = mu m c45 and the force pairs 0 < r < h:
//   a_i = C45 sum_j (A_i + A_j) F x_ij + Vc B_i sum_j (h - r) B_j (v_j - v_i) + g
// (Solver.cu:231-309,601-711 divided by rho_i), then v' = v + dt a,
// x' = x + dt v' and the wall clamp (Solver.cu:713-755).  The adjoint sweeps
// the steps backwards: clamp and integration per particle, then one pair pass
// for the adjoints of A, B, the force's position and velocity terms and the
// parameters, then one pair pass for the density's position term.  Pair sums
// are taken as symmetric (j in i's stencil iff i in j's).  Tension is not
// differentiated (the rollout is defined for tension 0).
// ---------------------------------------------------------------------------
namespace {

struct StepRecord {
    State s;                       // sorted pre-step state
    Neighbourhood nb;
    std::vector<double> rho, acc;
};

template <class F>
void for_each_pair(const oracle_params& p, const State& s, const Neighbourhood& nb, uint32_t i,
                   F&& f) {
    const double* xi = &s.x[3 * i];
    for_each_stencil_cell(p, xi, [&](uint32_t c) {
        const uint32_t b = nb.start[c], e = nb.end[c];
        if (b == kEmpty) return;
        for (uint32_t j = b; j < e; ++j) {
            const double xij[3] = {xi[0] - s.x[3 * j], xi[1] - s.x[3 * j + 1],
                                   xi[2] - s.x[3 * j + 2]};
            const double r = std::sqrt(xij[0] * xij[0] + xij[1] * xij[1] + xij[2] * xij[2]);
            if (r != 0.0 && r < p.h) f(j, xij, r);
        }
    });
}

// One step's adjoint: (gx, gv) hold dL/dx', dL/dv' in the step's sorted order
// on entry and dL/dx, dL/dv in the step's input order on exit.
void step_adjoint(const oracle_params& p, const StepRecord& R, double dt, std::vector<double>& gx,
                  std::vector<double>& gv, double& gk, double& gmu) {
    const uint32_t n = static_cast<uint32_t>(R.s.id.size());
    const Coeffs cf(p.h);
    const double C45 = p.mass * cf.c45, Vc = p.viscosity * p.mass * cf.c45;
    const double C6 = p.mass * cf.c6, h = p.h, h2 = p.h * p.h;
    std::vector<double> ga(3 * n), gxd(3 * n), gvd(3 * n), A(n), B(n), grho(n);
    // Clamp and symplectic Euler, per particle and axis.
    for (uint32_t i = 0; i < n; ++i) {
        A[i] = p.stiffness * (R.rho[i] - p.rest_density) / (R.rho[i] * R.rho[i]);
        B[i] = 1.0 / R.rho[i];
        for (int d = 0; d < 3; ++d) {
            const size_t k = 3 * static_cast<size_t>(i) + d;
            const double vp = R.s.v[k] + dt * R.acc[k];
            const double xp = R.s.x[k] + dt * vp;
            double gxp = gx[k], gvp = gv[k];
            if (p.wall_clamp && (xp < p.lo[d] || xp > p.hi[d])) {
                gxp = 0.0;
                gvp = -p.restitution * gv[k];
            }
            gvp += dt * gxp;
            gxd[k] = gxp;
            gvd[k] = gvp;
            ga[k] = dt * gvp;
        }
    }
    // Force pair pass.
    for (uint32_t i = 0; i < n; ++i) {
        const double* vi = &R.s.v[3 * i];
        const double* ai = &ga[3 * i];
        double sumA = 0.0, S1 = 0.0, S2 = 0.0, gxs[3] = {0, 0, 0}, gvs[3] = {0, 0, 0};
        for_each_pair(p, R.s, R.nb, i, [&](uint32_t j, const double* xij, double r) {
            const double* vj = &R.s.v[3 * j];
            const double* aj = &ga[3 * j];
            const double hr = h - r, F = hr * hr / r, Fr = -hr * (h + r) / (r * r * r);
            double dA[3], dv[3];
            for (int d = 0; d < 3; ++d) {
                dA[d] = ai[d] - aj[d];
                dv[d] = vj[d] - vi[d];
            }
            const double xdA = xij[0] * dA[0] + xij[1] * dA[1] + xij[2] * dA[2];
            const double dvdA = dv[0] * dA[0] + dv[1] * dA[1] + dv[2] * dA[2];
            sumA += F * xdA;
            S1 += hr * B[j] * (dv[0] * ai[0] + dv[1] * ai[1] + dv[2] * ai[2]);
            S2 += hr * B[j] * (dv[0] * aj[0] + dv[1] * aj[1] + dv[2] * aj[2]);
            const double pa = A[i] + A[j];
            const double cx = C45 * pa * Fr * xdA - Vc * B[i] * B[j] * dvdA / r;
            for (int d = 0; d < 3; ++d) {
                gxs[d] += C45 * pa * F * dA[d] + cx * xij[d];
                gvs[d] -= Vc * B[i] * B[j] * hr * dA[d];
            }
        });
        const double gA = C45 * sumA, gB = Vc * (S1 - S2);
        grho[i] = gA * p.stiffness * (2.0 * p.rest_density - R.rho[i]) * B[i] * B[i] * B[i] -
                  gB * B[i] * B[i];
        gk += gA * (R.rho[i] - p.rest_density) * B[i] * B[i];
        gmu += p.mass * cf.c45 * B[i] * S1;
        for (int d = 0; d < 3; ++d) {
            gxd[3 * i + d] += gxs[d];
            gvd[3 * i + d] += gvs[d];
        }
    }
    // Density pair pass, then back to the step's input order.
    for (uint32_t i = 0; i < n; ++i) {
        double gxs[3] = {0, 0, 0};
        for_each_pair(p, R.s, R.nb, i, [&](uint32_t j, const double* xij, double r) {
            const double a = h2 - r * r;
            const double c = -6.0 * C6 * a * a * (grho[i] + grho[j]);
            for (int d = 0; d < 3; ++d) gxs[d] += c * xij[d];
        });
        const uint32_t q = R.nb.perm[i];
        for (int d = 0; d < 3; ++d) {
            gx[3 * q + d] = gxd[3 * i + d] + gxs[d];
            gv[3 * q + d] = gvd[3 * i + d];
        }
    }
}

}  // namespace

extern "C" {

double oracle_rollout_loss_grad(const oracle_params* p, const double* pos0, const double* vel0,
                                uint32_t n, double dt, uint32_t n_steps, const double* target,
                                double* x_final, double* g_pos0, double* g_vel0,
                                double* g_stiffness, double* g_viscosity) {
    State s;
    s.x.assign(pos0, pos0 + 3 * static_cast<size_t>(n));
    s.v.assign(3 * static_cast<size_t>(n), 0.0);
    if (vel0) s.v.assign(vel0, vel0 + 3 * static_cast<size_t>(n));
    s.id.resize(n);
    std::iota(s.id.begin(), s.id.end(), 0u);
    const bool grads = g_pos0 || g_vel0 || g_stiffness || g_viscosity;
    std::vector<StepRecord> rec(grads ? n_steps : 0);
    Neighbourhood nb;
    std::vector<double> r, a;
    for (uint32_t t = 0; t < n_steps; ++t) {
        build_neighbourhood(*p, s, nb);
        densities(*p, s, nb, r);
        accelerations(*p, s, nb, r, a);
        if (grads) rec[t] = StepRecord{s, nb, r, a};
        integrate(*p, s, a, dt);
    }
    // L = sum (x_T - target)^2 / (3n), particles matched by id.
    double loss = 0.0;
    std::vector<double> gx(3 * static_cast<size_t>(n)), gv(3 * static_cast<size_t>(n), 0.0);
    const double scale = 1.0 / (3.0 * n);
    for (uint32_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            const double e = s.x[3 * i + d] - target[3 * s.id[i] + d];
            loss += e * e * scale;
            gx[3 * i + d] = 2.0 * e * scale;
            if (x_final) x_final[3 * s.id[i] + d] = s.x[3 * i + d];
        }
    if (!grads) return loss;
    double gk = 0.0, gmu = 0.0;
    for (uint32_t t = n_steps; t-- > 0;) step_adjoint(*p, rec[t], dt, gx, gv, gk, gmu);
    if (g_pos0) std::copy(gx.begin(), gx.end(), g_pos0);
    if (g_vel0) std::copy(gv.begin(), gv.end(), g_vel0);
    if (g_stiffness) *g_stiffness = gk;
    if (g_viscosity) *g_viscosity = gmu;
    return loss;
}

}  // extern "C"
