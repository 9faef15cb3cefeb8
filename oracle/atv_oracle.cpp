// atv_oracle.cpp — CPU FP64 restatement of the reference classical-MD path.
//
// TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg. Built with
// -ffp-contract=off so every expression rounds exactly as written, which is
// what the bit-exact pair-set comparison relies on.
//
// Citations are to the reference tree (ELOISE-Wang0214/Atomiverse).

#include "atv_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// include/atomiverse/core/units.h:138,165-166
constexpr double kB_ev_per_k = 8.617333262e-5;
constexpr double force_to_accel = 1.0 / 103.6427;
constexpr double mass_v2_to_energy = 103.6427;

struct Grid {
    int nc[3];
    double cs[3];
    double L[3];
    int reach[3];
    std::vector<std::array<int, 3>> half, full;
    double rc_sq;
};

// CellList::compute_grid (src/cpu/neighbor/cell_list.cpp:83-154, single-rank
// branch) + CellList::build_stencil (:179-232).
Grid make_grid(const double box[3], double cutoff_skin) {
    Grid g;
    for (int d = 0; d < 3; ++d) {
        const double L = box[d];
        const int nc = std::max(int(1), static_cast<int>(std::floor(L / (0.5 * cutoff_skin))));
        g.nc[d] = nc;
        g.cs[d] = L / nc;
        g.L[d] = L;
    }
    for (int d = 0; d < 3; ++d)
        g.reach[d] = std::max(1, static_cast<int>(std::ceil(cutoff_skin / g.cs[d] - 1e-12)));
    const double rc_sq = cutoff_skin * cutoff_skin;
    g.rc_sq = rc_sq;
    auto min_dist_sq = [&](int di, int dj, int dk) {
        const double ddx = di > 0 ? (di - 1) * g.cs[0] : (di < 0 ? (di + 1) * g.cs[0] : 0.0);
        const double ddy = dj > 0 ? (dj - 1) * g.cs[1] : (dj < 0 ? (dj + 1) * g.cs[1] : 0.0);
        const double ddz = dk > 0 ? (dk - 1) * g.cs[2] : (dk < 0 ? (dk + 1) * g.cs[2] : 0.0);
        return ddx * ddx + ddy * ddy + ddz * ddz;
    };
    for (int di = -g.reach[0]; di <= g.reach[0]; ++di)
        for (int dj = -g.reach[1]; dj <= g.reach[1]; ++dj)
            for (int dk = -g.reach[2]; dk <= g.reach[2]; ++dk) {
                if (min_dist_sq(di, dj, dk) > rc_sq)
                    continue;
                g.full.push_back({di, dj, dk});
                const bool home = (di == 0 && dj == 0 && dk == 0);
                const bool forward = (di > 0) || (di == 0 && dj > 0) || (di == 0 && dj == 0 && dk > 0);
                if (home || forward)
                    g.half.push_back({di, dj, dk});
            }
    return g;
}

struct Bins {
    std::vector<int> atom_cell, cell_count, cell_offset, binned;
};

// CellList::bin_atoms (cell_list.cpp:234-314), single-rank: wrap into [0, L)
// for the cell index only, floor((x - 0)/cs), clamp, z-fastest cell index,
// then a serial counting sort that keeps ascending i inside a cell.
void bin_atoms(const Grid& g, int64_t n, const double* x, const double* y, const double* z,
               Bins& b) {
    const int ncells = g.nc[0] * g.nc[1] * g.nc[2];
    b.atom_cell.assign(n, 0);
    b.cell_count.assign(ncells, 0);
    b.cell_offset.assign(ncells, 0);
    b.binned.assign(n, 0);
    const double lx = g.L[0], ly = g.L[1], lz = g.L[2];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double xi = x[i], yi = y[i], zi = z[i];
        xi -= lx * std::floor(xi / lx);
        yi -= ly * std::floor(yi / ly);
        zi -= lz * std::floor(zi / lz);
        int cx = static_cast<int>(std::floor((xi - 0.0) / g.cs[0]));
        int cy = static_cast<int>(std::floor((yi - 0.0) / g.cs[1]));
        int cz = static_cast<int>(std::floor((zi - 0.0) / g.cs[2]));
        cx = std::max(0, std::min(cx, g.nc[0] - 1));
        cy = std::max(0, std::min(cy, g.nc[1] - 1));
        cz = std::max(0, std::min(cz, g.nc[2] - 1));
        b.atom_cell[i] = cx * g.nc[1] * g.nc[2] + cy * g.nc[2] + cz;
    }
    for (int64_t i = 0; i < n; ++i)
        b.cell_count[b.atom_cell[i]]++;
    b.cell_offset[0] = 0;
    for (int c = 1; c < ncells; ++c)
        b.cell_offset[c] = b.cell_offset[c - 1] + b.cell_count[c - 1];
    std::vector<int> wp(b.cell_offset);
    for (int64_t i = 0; i < n; ++i) {
        const int c = b.atom_cell[i];
        b.binned[wp[c]++] = static_cast<int>(i);
    }
}

// One row of CellList::refresh_half, serial path (cell_list.cpp:539-738):
// half stencil with a per-cell +/-L shift when every axis has >= 2*reach+1
// cells, else the full stencil with j<=i rejection and while-loop min-image.
// Appends accepted j to `out`.
template <class Out>
inline void half_row(const Grid& g, const Bins& b, int64_t i, const double* x, const double* y,
                     const double* z, bool use_half, Out& out) {
    const auto& stencil = use_half ? g.half : g.full;
    const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2];
    const double lx = g.L[0], ly = g.L[1], lz = g.L[2];
    const double half_lx = 0.5 * lx, half_ly = 0.5 * ly, half_lz = 0.5 * lz;
    const int ci = b.atom_cell[i];
    const int cz0 = ci % ncz;
    const int cy0 = (ci / ncz) % ncy;
    const int cx0 = ci / (ncy * ncz);
    const double xi = x[i], yi = y[i], zi = z[i];
    int visited[128];
    int visited_n = 0;
    for (const auto& s : stencil) {
        int cx = cx0 + s[0], cy = cy0 + s[1], cz = cz0 + s[2];
        double shx = 0.0, shy = 0.0, shz = 0.0;
        if (cx < 0) { cx += ncx; shx = -lx; } else if (cx >= ncx) { cx -= ncx; shx = lx; }
        if (cx < 0 || cx >= ncx) cx = ((cx % ncx) + ncx) % ncx;
        if (cy < 0) { cy += ncy; shy = -ly; } else if (cy >= ncy) { cy -= ncy; shy = ly; }
        if (cy < 0 || cy >= ncy) cy = ((cy % ncy) + ncy) % ncy;
        if (cz < 0) { cz += ncz; shz = -lz; } else if (cz >= ncz) { cz -= ncz; shz = lz; }
        if (cz < 0 || cz >= ncz) cz = ((cz % ncz) + ncz) % ncz;
        const int cj = cx * ncy * ncz + cy * ncz + cz;
        bool seen = false;
        for (int v = 0; v < visited_n; ++v)
            if (visited[v] == cj) { seen = true; break; }
        if (seen)
            continue;
        visited[visited_n++] = cj;
        const int start = b.cell_offset[cj];
        const int end = start + b.cell_count[cj];
        const bool reject_le_i = !use_half || (cj == ci);
        for (int idx = start; idx < end; ++idx) {
            const int j = b.binned[idx];
            if (j <= i && reject_le_i)
                continue;
            double dx, dy, dz;
            if (use_half) {
                dx = xi - (x[j] + shx);
                dy = yi - (y[j] + shy);
                dz = zi - (z[j] + shz);
            } else {
                dx = xi - x[j];
                dy = yi - y[j];
                dz = zi - z[j];
                while (dx > half_lx) dx -= lx;
                while (dx < -half_lx) dx += lx;
                while (dy > half_ly) dy -= ly;
                while (dy < -half_ly) dy += ly;
                while (dz > half_lz) dz -= lz;
                while (dz < -half_lz) dz += lz;
            }
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < g.rc_sq)
                out.push_back(j);
        }
    }
}

bool grid_use_half(const Grid& g) {
    return g.nc[0] >= 2 * g.reach[0] + 1 && g.nc[1] >= 2 * g.reach[1] + 1 &&
           g.nc[2] >= 2 * g.reach[2] + 1;
}

struct HalfList {
    std::vector<int> numneigh;
    std::vector<int64_t> firstneigh;
    std::vector<int> neighbors;
};

// CellList::refresh_half rows computed in contiguous chunks per thread and
// concatenated in row order: the CSR equals the serial path's exactly.
void build_half(const Grid& g, const Bins& b, int64_t n, const double* x, const double* y,
                const double* z, int nthreads, HalfList& hl) {
    hl.numneigh.assign(n, 0);
    hl.firstneigh.assign(n, 0);
    const bool use_half = grid_use_half(g);
    if (nthreads < 1)
        nthreads = 1;
    std::vector<std::vector<int>> chunk(nthreads);
    std::vector<int64_t> lo(nthreads + 1);
    for (int t = 0; t <= nthreads; ++t)
        lo[t] = n * t / nthreads;
#pragma omp parallel for schedule(static, 1) num_threads(nthreads)
    for (int t = 0; t < nthreads; ++t) {
        auto& out = chunk[t];
        out.clear();
        out.reserve(static_cast<size_t>((lo[t + 1] - lo[t]) * 48));
        for (int64_t i = lo[t]; i < lo[t + 1]; ++i) {
            const size_t before = out.size();
            half_row(g, b, i, x, y, z, use_half, out);
            hl.numneigh[i] = static_cast<int>(out.size() - before);
        }
    }
    int64_t total = 0;
    std::vector<int64_t> base(nthreads);
    for (int t = 0; t < nthreads; ++t) {
        base[t] = total;
        total += static_cast<int64_t>(chunk[t].size());
    }
    hl.neighbors.resize(total);
#pragma omp parallel for schedule(static, 1) num_threads(nthreads)
    for (int t = 0; t < nthreads; ++t) {
        int64_t cur = base[t];
        for (int64_t i = lo[t]; i < lo[t + 1]; ++i) {
            hl.firstneigh[i] = cur;
            cur += hl.numneigh[i];
        }
        if (!chunk[t].empty())
            std::memcpy(hl.neighbors.data() + base[t], chunk[t].data(), chunk[t].size() * sizeof(int));
    }
}

inline void min_image_round(double& dx, double& dy, double& dz, const double L[3]) {
    // Box::min_image, orthorhombic (include/atomiverse/cpu/frame/box.h:123-140)
    dx -= L[0] * std::nearbyint(dx / L[0]);
    dy -= L[1] * std::nearbyint(dy / L[1]);
    dz -= L[2] * std::nearbyint(dz / L[2]);
}

// PairLJCut::compute (src/cpu/potentials/pair_lj_cut.cpp:300-830), scalar form
// with while-loop min-image (the has_shifts=false path, :730-749), per-thread
// force buffers reduced in thread order (:415-433, :806-817).
void pair_lj(int64_t n, const double* x, const double* y, const double* z, const double box[3],
             const int* numneigh, const int64_t* firstneigh, const int* neighbors, double eps,
             double sigma, double cutoff, int ntypes, const int* type, const double* eps_tab,
             const double* sigma_tab, int nthreads, double* fx, double* fy, double* fz,
             double* energy_out, double* virial_out) {
    const double lx = box[0], ly = box[1], lz = box[2];
    const double half_lx = 0.5 * lx, half_ly = 0.5 * ly, half_lz = 0.5 * lz;
    const double rc_sq = cutoff * cutoff;
    // constructor (pair_lj_cut.cpp:31-42)
    const double sig6 = sigma * sigma * sigma * sigma * sigma * sigma;
    double vshift = 0.0;
    {
        const double r2inv = 1.0 / rc_sq;
        const double r6inv = r2inv * r2inv * r2inv * sig6;
        vshift = 4.0 * eps * (r6inv * r6inv - r6inv);
    }
    const double eps4 = 4.0 * eps, eps24 = 24.0 * eps;
    // finalize_type_table_ (pair_lj_cut.cpp:257-298): sigma^6 and shift per pair
    std::vector<double> s6tab, shtab;
    if (ntypes > 0) {
        s6tab.resize(ntypes * ntypes);
        shtab.resize(ntypes * ntypes);
        for (int ab = 0; ab < ntypes * ntypes; ++ab) {
            const double s = sigma_tab[ab], e = eps_tab[ab];
            const double s2 = s * s;
            const double s6 = s2 * s2 * s2;
            s6tab[ab] = s6;
            const double r2inv = 1.0 / (cutoff * cutoff);
            const double r6inv = r2inv * r2inv * r2inv * s6;
            shtab[ab] = 4.0 * e * (r6inv * r6inv - r6inv);
        }
    }
    if (nthreads < 1)
        nthreads = 1;
    std::vector<std::vector<double>> buf(nthreads > 1 ? nthreads : 0);
    double e_tot = 0, v_xx = 0, v_yy = 0, v_zz = 0, v_xy = 0, v_xz = 0, v_yz = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : e_tot, v_xx, v_yy, v_zz, v_xy, v_xz, v_yz)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *fxt, *fyt, *fzt;
        if (nthreads == 1) {
            fxt = fx; fyt = fy; fzt = fz;
        } else {
            buf[tid].assign(3 * n, 0.0);
            fxt = buf[tid].data();
            fyt = fxt + n;
            fzt = fyt + n;
        }
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const int jnum = numneigh[i];
            const int64_t off = firstneigh[i];
            const double xi = x[i], yi = y[i], zi = z[i];
            double fi_x = 0, fi_y = 0, fi_z = 0;
            for (int jj = 0; jj < jnum; ++jj) {
                const int j = neighbors[off + jj];
                double dx = xi - x[j];
                double dy = yi - y[j];
                double dz = zi - z[j];
                while (dx > half_lx) dx -= lx;
                while (dx < -half_lx) dx += lx;
                while (dy > half_ly) dy -= ly;
                while (dy < -half_ly) dy += ly;
                while (dz > half_lz) dz -= lz;
                while (dz < -half_lz) dz += lz;
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 >= rc_sq)
                    continue;
                double p_sig6 = sig6, p_eps24 = eps24, p_eps4 = eps4, p_shift = vshift;
                if (ntypes > 0) {
                    const int ab = type[i] * ntypes + type[j];
                    p_sig6 = s6tab[ab];
                    const double e = eps_tab[ab];
                    p_eps24 = 24.0 * e;
                    p_eps4 = 4.0 * e;
                    p_shift = shtab[ab];
                }
                const double r2inv = 1.0 / r2;
                const double r6inv = r2inv * r2inv * r2inv * p_sig6;
                const double fpair = p_eps24 * r2inv * (2.0 * r6inv * r6inv - r6inv);
                const double fxp = fpair * dx, fyp = fpair * dy, fzp = fpair * dz;
                fi_x += fxp;
                fi_y += fyp;
                fi_z += fzp;
                fxt[j] -= fxp;
                fyt[j] -= fyp;
                fzt[j] -= fzp;
                v_xx += dx * fxp;
                v_yy += dy * fyp;
                v_zz += dz * fzp;
                v_xy += dx * fyp;
                v_xz += dx * fzp;
                v_yz += dy * fzp;
                e_tot += p_eps4 * (r6inv * r6inv - r6inv) - p_shift;
            }
            fxt[i] += fi_x;
            fyt[i] += fi_y;
            fzt[i] += fi_z;
        }
        if (nthreads > 1) {
#pragma omp barrier
#pragma omp for schedule(static)
            for (int64_t i = 0; i < n; ++i)
                for (int t = 0; t < nthreads; ++t) {
                    fx[i] += buf[t][i];
                    fy[i] += buf[t][n + i];
                    fz[i] += buf[t][2 * n + i];
                }
        }
    }
    if (energy_out)
        *energy_out += e_tot;
    if (virial_out) {
        virial_out[0] += v_xx;
        virial_out[1] += v_yy;
        virial_out[2] += v_zz;
        virial_out[3] += v_xy;
        virial_out[4] += v_xz;
        virial_out[5] += v_yz;
    }
}

// KineticEnergy::launch (src/cpu/dynamics/init.cpp:29-46)
double kinetic(int64_t n, const double* mass, const double* vx, const double* vy,
               const double* vz) {
    double ke = 0;
#pragma omp parallel for schedule(static) reduction(+ : ke)
    for (int64_t i = 0; i < n; ++i) {
        const double v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
        ke += mass[i] * v2;
    }
    return 0.5 * mass_v2_to_energy * ke;
}

} // namespace

extern "C" {

int64_t orc_fcc_lattice(int nx, int ny, int nz, double a, double* x, double* y, double* z) {
    static const double basis[4][3] = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};
    int64_t k = 0;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix)
                for (int b = 0; b < 4; ++b) {
                    if (x) {
                        x[k] = (ix + basis[b][0]) * a;
                        y[k] = (iy + basis[b][1]) * a;
                        z[k] = (iz + basis[b][2]) * a;
                    }
                    ++k;
                }
    return k;
}

// InitVelocity::build, gaussian branch (src/cpu/dynamics/init.cpp:107-215)
void orc_init_velocity(int64_t n, const double* mass, double T, uint32_t seed, int64_t ndof,
                       double* vx, double* vy, double* vz) {
    std::mt19937 rng(seed);
    std::normal_distribution<double> draw(0.0, 1.0);
    for (int64_t i = 0; i < n; ++i) {
        const double s = 1.0 / std::sqrt(mass[i]);
        vx[i] = draw(rng) * s;
        vy[i] = draw(rng) * s;
        vz[i] = draw(rng) * s;
    }
    double vcm_x = 0, vcm_y = 0, vcm_z = 0, mtot = 0;
    for (int64_t i = 0; i < n; ++i) {
        vcm_x += mass[i] * vx[i];
        vcm_y += mass[i] * vy[i];
        vcm_z += mass[i] * vz[i];
        mtot += mass[i];
    }
    if (mtot > 0) {
        vcm_x /= mtot;
        vcm_y /= mtot;
        vcm_z /= mtot;
        for (int64_t i = 0; i < n; ++i) {
            vx[i] -= vcm_x;
            vy[i] -= vcm_y;
            vz[i] -= vcm_z;
        }
    }
    double sum_mv2 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
        sum_mv2 += mass[i] * v2;
    }
    if (ndof > 0 && sum_mv2 > 0.0) {
        const double ke = 0.5 * mass_v2_to_energy * sum_mv2;
        const double t_measured = 2.0 * ke / (static_cast<double>(ndof) * kB_ev_per_k);
        const double factor = std::sqrt(T / t_measured);
        for (int64_t i = 0; i < n; ++i) {
            vx[i] *= factor;
            vy[i] *= factor;
            vz[i] *= factor;
        }
    }
}

int orc_cell_grid(const double box[3], double cutoff_skin, int* out_nc, double* out_cs,
                  int* out_reach, int* out_nfull) {
    Grid g = make_grid(box, cutoff_skin);
    for (int d = 0; d < 3; ++d) {
        out_nc[d] = g.nc[d];
        out_cs[d] = g.cs[d];
        out_reach[d] = g.reach[d];
    }
    if (out_nfull)
        *out_nfull = static_cast<int>(g.full.size());
    return static_cast<int>(g.half.size());
}

int orc_bin_atoms(int64_t n, const double* x, const double* y, const double* z, const double box[3],
                  double cutoff_skin, int* atom_cell, int* cell_count, int* cell_offset, int* binned) {
    Grid g = make_grid(box, cutoff_skin);
    Bins b;
    bin_atoms(g, n, x, y, z, b);
    std::memcpy(atom_cell, b.atom_cell.data(), n * sizeof(int));
    std::memcpy(binned, b.binned.data(), n * sizeof(int));
    std::memcpy(cell_count, b.cell_count.data(), b.cell_count.size() * sizeof(int));
    std::memcpy(cell_offset, b.cell_offset.data(), b.cell_offset.size() * sizeof(int));
    return static_cast<int>(b.cell_count.size());
}

int64_t orc_build_half(int64_t n, const double* x, const double* y, const double* z,
                       const double box[3], double cutoff_skin, int nthreads, int* numneigh,
                       int64_t* firstneigh, int* neighbors, int64_t cap) {
    Grid g = make_grid(box, cutoff_skin);
    Bins b;
    bin_atoms(g, n, x, y, z, b);
    HalfList hl;
    build_half(g, b, n, x, y, z, nthreads, hl);
    const int64_t total = static_cast<int64_t>(hl.neighbors.size());
    if (total > cap)
        return -total;
    std::memcpy(numneigh, hl.numneigh.data(), n * sizeof(int));
    std::memcpy(firstneigh, hl.firstneigh.data(), n * sizeof(int64_t));
    if (total)
        std::memcpy(neighbors, hl.neighbors.data(), total * sizeof(int));
    return total;
}

// CellList::needs_rebuild (cell_list.cpp:355-392), displacement part
int orc_needs_rebuild(int64_t n, const double* x, const double* y, const double* z,
                      const double* xs, const double* ys, const double* zs, const double box[3],
                      double skin, double* max_disp_sq) {
    double m = 0;
    for (int64_t i = 0; i < n; ++i) {
        double dx = x[i] - xs[i], dy = y[i] - ys[i], dz = z[i] - zs[i];
        min_image_round(dx, dy, dz, box);
        const double dsq = dx * dx + dy * dy + dz * dz;
        if (dsq > m)
            m = dsq;
    }
    if (max_disp_sq)
        *max_disp_sq = m;
    return m * 4.0 > skin * skin ? 1 : 0;
}

void orc_pair_lj(int64_t n, const double* x, const double* y, const double* z, const double box[3],
                 const int* numneigh, const int64_t* firstneigh, const int* neighbors, double eps,
                 double sigma, double cutoff, int ntypes, const int* type, const double* eps_tab,
                 const double* sigma_tab, int nthreads, double* fx, double* fy, double* fz,
                 double* energy, double* virial) {
    pair_lj(n, x, y, z, box, numneigh, firstneigh, neighbors, eps, sigma, cutoff, ntypes, type,
            eps_tab, sigma_tab, nthreads, fx, fy, fz, energy, virial);
}

// lorentz_berthelot (src/cpu/potentials/pair_lj_cut.cpp:245-252): eps_ij =
// sqrt(eps_i eps_j), sigma_ij = (sigma_i + sigma_j)/2.
void orc_lj_mix(int ntypes, double* eps_tab, double* sigma_tab) {
    for (int a = 0; a < ntypes; ++a)
        for (int b = 0; b < ntypes; ++b) {
            if (a == b)
                continue;
            const double eii = eps_tab[a * ntypes + a], ejj = eps_tab[b * ntypes + b];
            const double sii = sigma_tab[a * ntypes + a], sjj = sigma_tab[b * ntypes + b];
            eps_tab[a * ntypes + b] = std::sqrt(eii * ejj);
            sigma_tab[a * ntypes + b] = 0.5 * (sii + sjj);
        }
}

double orc_kinetic_energy(int64_t n, const double* mass, const double* vx, const double* vy,
                          const double* vz) {
    return kinetic(n, mass, vx, vy, vz);
}

// cpu::md::Virial::launch (src/cpu/observe/virial_stress.cpp:14-56): the
// kinetic stress, serial sums in atom order, each term m * v_a * v_b as the
// reference writes it, scaled by mass_v2_to_energy.
void orc_kinetic_stress(int64_t n, const double* mass, const double* vx, const double* vy, const double* vz,
                        double out[6]) {
    double v_xx = 0, v_yy = 0, v_zz = 0, v_xy = 0, v_xz = 0, v_yz = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double m = mass[i];
        v_xx += m * vx[i] * vx[i];
        v_yy += m * vy[i] * vy[i];
        v_zz += m * vz[i] * vz[i];
        v_xy += m * vx[i] * vy[i];
        v_xz += m * vx[i] * vz[i];
        v_yz += m * vy[i] * vz[i];
    }
    out[0] = mass_v2_to_energy * v_xx;
    out[1] = mass_v2_to_energy * v_yy;
    out[2] = mass_v2_to_energy * v_zz;
    out[3] = mass_v2_to_energy * v_xy;
    out[4] = mass_v2_to_energy * v_xz;
    out[5] = mass_v2_to_energy * v_yz;
}

// cpu::md::MD::run driving Runtime::step_verlet_ (src/cpu/driver/runtime.cpp:452-502):
//   initial_integrate: VelocityVerlet (velocity_verlet.cpp:41-55) then
//                      NoseHoover (nose_hoover.cpp:104-170)
//   decide: NeighborList::decide_reneighbor (neighbor_list.h:128-141)
//   neighbor: CellList::build + refresh_half when decided
//   pre_force: ZeroForces / ZeroEnergy;  force: PairLJCut::compute
//   final_integrate: VelocityVerlet then NoseHoover
//   end_of_step: KineticEnergy
int orc_md_run(int64_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
               double* fx, double* fy, double* fz, const double* mass, const double box[3],
               const orc_md_params* p, orc_md_state* st, int nsteps, double* pe, double* ke,
               double* virial_out) {
    return orc_md_run_ff(n, x, y, z, vx, vy, vz, fx, fy, fz, mass, box, p, nullptr, st, nsteps, pe,
                         ke, virial_out);
}

// Same loop; the force phase is PairLJCut, or PairLJCutCoulLong followed by
// KSpacePME when `ff` is given (evaluators run in insertion order at
// Phase::force, include/atomiverse/core/plan.h stable sort).
int orc_md_run_ff(int64_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                  double* fx, double* fy, double* fz, const double* mass, const double box[3],
                  const orc_md_params* p, const orc_md_ff* ff, orc_md_state* st, int nsteps,
                  double* pe, double* ke, double* virial_out) {
    if (n <= 0 || !p || !st)
        return -1;
    const int nthreads = p->nthreads < 1 ? 1 : p->nthreads;
#ifdef _OPENMP
    omp_set_num_threads(nthreads);
#endif
    const double cutoff_skin = p->cutoff + p->skin;
    Grid g = make_grid(box, cutoff_skin);
    Bins b;
    HalfList hl;
    std::vector<double> xs(n), ys(n), zs(n);
    auto rebuild = [&]() {
        bin_atoms(g, n, x, y, z, b);
        std::memcpy(xs.data(), x, n * sizeof(double));
        std::memcpy(ys.data(), y, n * sizeof(double));
        std::memcpy(zs.data(), z, n * sizeof(double));
        build_half(g, b, n, x, y, z, nthreads, hl);
    };
    // NeighborList::build at component build (neighbor_list.h:146-171); the list
    // is not persisted between calls, so a continued run re-enumerates it from
    // the current positions (same pair set a resident list would give only if
    // no rebuild was due; tests that continue runs rebuild at the boundary).
    rebuild();
    if (st->step == 0) {
        st->ago = 0;
        st->nrebuilds = 1;
    }
    const int64_t ndof = p->ndof > 0 ? p->ndof : 3 * n - 3;
    const double Q = static_cast<double>(ndof) * kB_ev_per_k * p->T_target * p->tau_T * p->tau_T;
    const double dt = p->dt;
    const double dtf = 0.5 * dt * force_to_accel;
    const double lx = box[0], ly = box[1], lz = box[2];
    auto nose_hoover = [&]() {
        const double k = kinetic(n, mass, vx, vy, vz);
        st->xi += (dt * 0.5) * (2.0 * k - static_cast<double>(ndof) * kB_ev_per_k * p->T_target) / Q;
        const double scale = std::exp(-dt * 0.5 * st->xi);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            vx[i] *= scale;
            vy[i] *= scale;
            vz[i] *= scale;
        }
    };
    for (int s = 0; s < nsteps; ++s) {
        // initial_integrate
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const double dtfm = dtf / mass[i];
            vx[i] += dtfm * fx[i];
            vy[i] += dtfm * fy[i];
            vz[i] += dtfm * fz[i];
            x[i] += dt * vx[i];
            y[i] += dt * vy[i];
            z[i] += dt * vz[i];
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            x[i] -= lx * std::floor(x[i] / lx);
            y[i] -= ly * std::floor(y[i] / ly);
            z[i] -= lz * std::floor(z[i] / lz);
        }
        if (p->nh)
            nose_hoover();
        // decide
        st->ago++;
        const bool delay_ok = st->ago >= p->delay;
        const bool every_ok = (st->ago % (p->every < 1 ? 1 : p->every)) == 0;
        bool nflag = false;
        if (delay_ok && every_ok) {
            nflag = !p->check || orc_needs_rebuild(n, x, y, z, xs.data(), ys.data(), zs.data(), box,
                                                   p->skin, nullptr);
        }
        if (nflag) {
            rebuild();
            st->ago = 0;
            st->nrebuilds++;
        }
        // pre_force + force
        std::memset(fx, 0, n * sizeof(double));
        std::memset(fy, 0, n * sizeof(double));
        std::memset(fz, 0, n * sizeof(double));
        double e = 0.0, vir[6] = {0, 0, 0, 0, 0, 0};
        if (!ff) {
            pair_lj(n, x, y, z, box, hl.numneigh.data(), hl.firstneigh.data(), hl.neighbors.data(),
                    p->eps, p->sigma, p->cutoff, 0, nullptr, nullptr, nullptr, nthreads, fx, fy, fz,
                    &e, vir);
        } else {
            orc_pair_lj_coul_long(n, x, y, z, ff->q, ff->type, box, hl.numneigh.data(),
                                  hl.firstneigh.data(), hl.neighbors.data(), ff->ntypes, ff->lj1,
                                  ff->lj2, ff->lj3, ff->lj4, ff->shift, p->cutoff, ff->e_tab,
                                  ff->f_tab, ff->nbin, nthreads, fx, fy, fz, &e, vir);
            if (ff->pme_order > 0)
                orc_pme(n, x, y, z, ff->q, box, ff->alpha, ff->pme_K, ff->pme_order, fx, fy, fz,
                        &e, vir, nullptr);
        }
        // final_integrate
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const double dtfm = dtf / mass[i];
            vx[i] += dtfm * fx[i];
            vy[i] += dtfm * fy[i];
            vz[i] += dtfm * fz[i];
        }
        if (p->nh)
            nose_hoover();
        // end_of_step
        if (pe)
            pe[s] = e;
        if (ke)
            ke[s] = kinetic(n, mass, vx, vy, vz);
        if (virial_out && s == nsteps - 1)
            std::memcpy(virial_out, vir, sizeof(vir));
        st->step++;
    }
    return 0;
}

} // extern "C"

// ---- MTK NoseHooverNPT ------------------------------------------------------
//
// cpu::md::NoseHooverNPT (src/cpu/dynamics/nose_hoover_npt.cpp), without the
// couple-xy / surface-tension options (the CUDA sibling,
// include/atomiverse/cuda/dynamics/nose_hoover_npt.cuh, has neither). Every
// expression keeps the reference's operation order; reductions are serial.

namespace {

// nose_hoover_npt.cpp:45
constexpr double kEpsDtMaxFraction = 0.005;

// expm1c (nose_hoover_npt.cpp:66-73): (e^{2x} - 1)/(2x), Taylor branch near 0
double npt_expm1c(double x) {
    if (std::abs(x) < 1e-6)
        return 1.0 + x + (2.0 / 3.0) * x * x;
    return std::expm1(2.0 * x) / (2.0 * x);
}

struct Mtk {
    const orc_npt_params* q;
    orc_npt_state* s;
    int64_t n;
    const double* mass;
    double *vx, *vy, *vz;
    double dt;
    int64_t ndof;
    int pdim;
    double kT, W;
    double Q_t[ORC_NPT_MAXCHAIN], Q_b[ORC_NPT_MAXCHAIN];
    double tdrag, pdrag;

    // compute_global_KE (:205-233)
    double global_ke() const {
        double ke = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            const double v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
            ke += mass[i] * v2;
        }
        ke *= 0.5 * mass_v2_to_energy;
        return ke;
    }
    // compute_axis_KE (:236-266)
    void axis_ke(double out[3]) const {
        double kx = 0.0, ky = 0.0, kz = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            const double m = mass[i];
            kx += m * vx[i] * vx[i];
            ky += m * vy[i] * vy[i];
            kz += m * vz[i] * vz[i];
        }
        out[0] = 0.5 * mass_v2_to_energy * kx;
        out[1] = 0.5 * mass_v2_to_energy * ky;
        out[2] = 0.5 * mass_v2_to_energy * kz;
    }
    // compute_temp_target (:194-203) and the per-step Q_t refresh (:838-842)
    void temp_target(int64_t step) {
        const double T_stop = q->T_stop < 0.0 ? q->T_target : q->T_stop;
        double delta = 0.0;
        if (q->nlast > q->nfirst) {
            delta = static_cast<double>(step - q->nfirst) / static_cast<double>(q->nlast - q->nfirst);
            delta = std::clamp(delta, 0.0, 1.0);
        }
        const double T_live = q->T_target + delta * (T_stop - q->T_target);
        kT = kB_ev_per_k * T_live;
        const double tau_T2 = q->tau_T * q->tau_T;
        Q_t[0] = static_cast<double>(ndof) * kT * tau_T2;
        for (int k = 1; k < q->mtchain; ++k)
            Q_t[k] = kT * tau_T2;
    }
    // nhc_atoms_half (:269-340)
    void nhc_atoms_half(double half_dt) {
        const double quarter_dt = 0.5 * half_dt;
        const double eighth_dt = 0.5 * quarter_dt;
        const int M = std::min(q->mtchain, ORC_NPT_MAXCHAIN);
        double* ed = s->eta_t_dot;
        auto G = [&](int k) -> double {
            if (k == 0)
                return (2.0 * s->ke_global - static_cast<double>(ndof) * kT) / Q_t[0];
            return (Q_t[k - 1] * ed[k - 1] * ed[k - 1] - kT) / Q_t[k];
        };
        for (int k = M - 1; k >= 0; --k) {
            const double expfac = (k < M - 1) ? std::exp(-ed[k + 1] * eighth_dt) : 1.0;
            ed[k] *= expfac;
            ed[k] += G(k) * quarter_dt;
            ed[k] *= tdrag;
            ed[k] *= expfac;
        }
        const double scale = std::exp(-ed[0] * half_dt);
        for (int64_t i = 0; i < n; ++i) {
            vx[i] *= scale;
            vy[i] *= scale;
            vz[i] *= scale;
        }
        s->ke_global *= scale * scale;
        for (int k = 0; k < M; ++k)
            s->eta_t[k] += ed[k] * half_dt;
        for (int k = 0; k < M; ++k) {
            const double expfac = (k < M - 1) ? std::exp(-ed[k + 1] * eighth_dt) : 1.0;
            ed[k] *= expfac;
            ed[k] += G(k) * quarter_dt;
            ed[k] *= expfac;
        }
    }
    // nhc_baro_half (:346-421)
    void nhc_baro_half(double half_dt) {
        const double quarter_dt = 0.5 * half_dt;
        const double eighth_dt = 0.5 * quarter_dt;
        const int M = std::min(q->mpchain, ORC_NPT_MAXCHAIN);
        const double W_axis = W / 3.0;
        double* ed = s->eta_b_dot;
        auto baro_2ke = [&]() -> double {
            double t = 0.0;
            for (int a = 0; a < 3; ++a)
                if (q->couple[a])
                    t += W_axis * s->eps_dot[a] * s->eps_dot[a];
            return t;
        };
        auto G = [&](int k) -> double {
            if (k == 0)
                return (baro_2ke() - kT) / Q_b[0];
            return (Q_b[k - 1] * ed[k - 1] * ed[k - 1] - kT) / Q_b[k];
        };
        for (int k = M - 1; k >= 0; --k) {
            const double expfac = (k < M - 1) ? std::exp(-ed[k + 1] * eighth_dt) : 1.0;
            ed[k] *= expfac;
            ed[k] += G(k) * quarter_dt;
            ed[k] *= pdrag;
            ed[k] *= expfac;
        }
        const double baro_scale = std::exp(-ed[0] * half_dt);
        for (int a = 0; a < 3; ++a)
            if (q->couple[a])
                s->eps_dot[a] *= baro_scale;
        for (int k = 0; k < M; ++k)
            s->eta_b[k] += ed[k] * half_dt;
        for (int k = 0; k < M; ++k) {
            const double expfac = (k < M - 1) ? std::exp(-ed[k + 1] * eighth_dt) : 1.0;
            ed[k] *= expfac;
            ed[k] += G(k) * quarter_dt;
            ed[k] *= expfac;
        }
    }
    // update_eps_dot (:459-581), all-axes / per-axis branches (no couple xy)
    void update_eps_dot(const double box[3], double half_dt) {
        const double V = box[0] * box[1] * box[2];
        if (V <= 0.0)
            return;
        double ke_axis[3];
        axis_ke(ke_axis);
        double diag[3] = {s->vir[0] + 2.0 * ke_axis[0], s->vir[1] + 2.0 * ke_axis[1],
                          s->vir[2] + 2.0 * ke_axis[2]};
        if (pdim == 3) {
            const double ave = (1.0 / 3.0) * (diag[0] + diag[1] + diag[2]);
            diag[0] = diag[1] = diag[2] = ave;
        }
        double mtk_active_ke = 0.0;
        for (int a = 0; a < 3; ++a)
            if (q->couple[a])
                mtk_active_ke += ke_axis[a];
        const int64_t N_atoms = (ndof + 3) / 3;
        const double mtk_term =
            (pdim > 0) ? 6.0 * mtk_active_ke / (static_cast<double>(pdim) * N_atoms) : 0.0;
        for (int a = 0; a < 3; ++a) {
            if (!q->couple[a])
                continue;
            const double P_axis = diag[a] / V;
            const double G_eps_a = (3.0 * V * (P_axis - q->P_target) + mtk_term) / W;
            s->eps_dot[a] += G_eps_a * half_dt;
        }
    }
    // scale_v_by_baro (:591-637)
    void scale_v_by_baro(double half_dt) {
        double sum_eps_dot = 0.0;
        for (int a = 0; a < 3; ++a)
            if (q->couple[a])
                sum_eps_dot += s->eps_dot[a];
        const int64_t N_atoms = (ndof + 3) / 3;
        const double mtk_term2 = (pdim > 0) ? sum_eps_dot / (static_cast<double>(pdim) * N_atoms) : 0.0;
        const double sx = std::exp(-(s->eps_dot[0] + mtk_term2) * half_dt);
        const double sy = std::exp(-(s->eps_dot[1] + mtk_term2) * half_dt);
        const double sz = std::exp(-(s->eps_dot[2] + mtk_term2) * half_dt);
        for (int64_t i = 0; i < n; ++i) {
            vx[i] *= sx;
            vy[i] *= sy;
            vz[i] *= sz;
        }
        if (sx == sy && sy == sz)
            s->ke_global *= sx * sx;
        else
            s->ke_global = global_ke();
    }
    // kick_v_by_force (:639-660)
    void kick(const double* fx, const double* fy, const double* fz, double half_dt) {
        const double dtf = half_dt * force_to_accel;
        for (int64_t i = 0; i < n; ++i) {
            const double dtfm = dtf / mass[i];
            vx[i] += dtfm * fx[i];
            vy[i] += dtfm * fy[i];
            vz[i] += dtfm * fz[i];
        }
    }
    // propagate_x_and_box (:675-780), single rank: no ghosts
    void propagate(double* x, double* y, double* z, double box[3]) {
        double a[3], drift[3];
        for (int d = 0; d < 3; ++d) {
            const double eps_dt = std::clamp(s->eps_dot[d] * dt, -kEpsDtMaxFraction, kEpsDtMaxFraction);
            a[d] = std::exp(eps_dt);
            drift[d] = npt_expm1c(eps_dt * 0.5);
        }
        for (int64_t i = 0; i < n; ++i) {
            x[i] = x[i] * a[0] + vx[i] * dt * drift[0];
            y[i] = y[i] * a[1] + vy[i] * dt * drift[1];
            z[i] = z[i] * a[2] + vz[i] * dt * drift[2];
        }
        for (int d = 0; d < 3; ++d)
            box[d] *= a[d];
        // wrap_orthorhombic (include/atomiverse/cpu/frame/pbc.h:37-51) into the new box
        const double lx = box[0], ly = box[1], lz = box[2];
        for (int64_t i = 0; i < n; ++i) {
            x[i] -= lx * std::floor(x[i] / lx);
            y[i] -= ly * std::floor(y[i] / ly);
            z[i] -= lz * std::floor(z[i] / lz);
        }
    }
    // extended_hamiltonian (:785-820)
    double hamiltonian(double kinetic_e, double potential_e, const double box[3]) const {
        double H = kinetic_e + potential_e;
        const double W_axis = W / 3.0;
        for (int a = 0; a < 3; ++a)
            if (q->couple[a])
                H += 0.5 * W_axis * s->eps_dot[a] * s->eps_dot[a];
        for (int k = 0; k < q->mtchain; ++k)
            H += 0.5 * Q_t[k] * s->eta_t_dot[k] * s->eta_t_dot[k];
        for (int k = 0; k < q->mpchain; ++k)
            H += 0.5 * Q_b[k] * s->eta_b_dot[k] * s->eta_b_dot[k];
        H += static_cast<double>(ndof) * kT * s->eta_t[0];
        for (int k = 1; k < q->mtchain; ++k)
            H += kT * s->eta_t[k];
        for (int k = 0; k < q->mpchain; ++k)
            H += kT * s->eta_b[k];
        H += q->P_target * (box[0] * box[1] * box[2]);
        return H;
    }
};

} // namespace

extern "C" {

// Runtime::step_verlet_ (src/cpu/driver/runtime.cpp:452-502) with the NPT
// integrator's three bindings (nose_hoover_npt.h:259-263): initial_integrate
// P1-P6, final_integrate P7, end_of_step P8-P11 (NoseHooverNPT::launch,
// nose_hoover_npt.cpp:826-889). The list is rebuilt on the decide cadence
// when the box changed since the last build (CellList::needs_rebuild ->
// grid_stale, src/cpu/neighbor/cell_list.cpp:156-159, 369-370) or an atom
// moved more than half the skin, on a grid recomputed for the new box
// (cell_list.cpp:344-347).
int orc_npt_run(int64_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                double* fx, double* fy, double* fz, const double* mass, double box[3],
                const orc_md_params* p, const orc_npt_params* q, orc_md_state* st,
                orc_npt_state* ns, int nsteps, double* pe, double* ke, double* H, double* V) {
    if (n <= 0 || !p || !q || !st || !ns)
        return -1;
    if (q->mtchain < 1 || q->mtchain > ORC_NPT_MAXCHAIN || q->mpchain < 1 || q->mpchain > ORC_NPT_MAXCHAIN)
        return -1;
    const int nthreads = p->nthreads < 1 ? 1 : p->nthreads;
#ifdef _OPENMP
    omp_set_num_threads(nthreads);
#endif
    const double cutoff_skin = p->cutoff + p->skin;
    Grid g = make_grid(box, cutoff_skin);
    Bins b;
    HalfList hl;
    std::vector<double> xs(n), ys(n), zs(n);
    double box_built[3] = {box[0], box[1], box[2]};
    auto rebuild = [&]() {
        if (box[0] != box_built[0] || box[1] != box_built[1] || box[2] != box_built[2])
            g = make_grid(box, cutoff_skin);
        bin_atoms(g, n, x, y, z, b);
        std::memcpy(xs.data(), x, n * sizeof(double));
        std::memcpy(ys.data(), y, n * sizeof(double));
        std::memcpy(zs.data(), z, n * sizeof(double));
        build_half(g, b, n, x, y, z, nthreads, hl);
        for (int d = 0; d < 3; ++d)
            box_built[d] = box[d];
    };
    rebuild();
    if (st->step == 0) {
        st->ago = 0;
        st->nrebuilds = 1;
    }
    Mtk m{};
    m.q = q;
    m.s = ns;
    m.n = n;
    m.mass = mass;
    m.vx = vx;
    m.vy = vy;
    m.vz = vz;
    m.dt = p->dt;
    m.ndof = p->ndof > 0 ? p->ndof : 3 * n - 3;
    m.pdim = (q->couple[0] ? 1 : 0) + (q->couple[1] ? 1 : 0) + (q->couple[2] ? 1 : 0);
    // NoseHooverNPT::build (:101-190)
    m.kT = kB_ev_per_k * q->T_target;
    const double tau_T2 = q->tau_T * q->tau_T;
    const double tau_P2 = q->tau_P * q->tau_P;
    m.Q_t[0] = static_cast<double>(m.ndof) * m.kT * tau_T2;
    for (int k = 1; k < ORC_NPT_MAXCHAIN; ++k)
        m.Q_t[k] = m.kT * tau_T2;
    m.Q_b[0] = m.kT * tau_P2;
    for (int k = 1; k < ORC_NPT_MAXCHAIN; ++k)
        m.Q_b[k] = m.kT * tau_P2;
    m.W = (static_cast<double>(m.ndof) + 3.0) * m.kT * tau_P2;
    m.tdrag = 1.0 - q->drag * p->dt / q->tau_T;
    m.pdrag = 1.0 - q->drag * p->dt / q->tau_P;
    if (!ns->built) {
        for (int k = 0; k < ORC_NPT_MAXCHAIN; ++k)
            ns->eta_t[k] = ns->eta_t_dot[k] = ns->eta_b[k] = ns->eta_b_dot[k] = 0.0;
        for (int a = 0; a < 3; ++a)
            ns->eps_dot[a] = 0.0;
        ns->ke_global = 0.5 * static_cast<double>(m.ndof) * m.kT;
        ns->built = 1;
    }
    const double half_dt = 0.5 * p->dt;
    for (int s = 0; s < nsteps; ++s) {
        // initial_integrate: P1-P6
        m.temp_target(st->step);
        m.nhc_baro_half(half_dt);
        m.nhc_atoms_half(half_dt);
        m.update_eps_dot(box, half_dt);
        m.scale_v_by_baro(half_dt);
        m.kick(fx, fy, fz, half_dt);
        m.propagate(x, y, z, box);
        // decide
        st->ago++;
        const bool delay_ok = st->ago >= p->delay;
        const bool every_ok = (st->ago % (p->every < 1 ? 1 : p->every)) == 0;
        bool nflag = false;
        if (delay_ok && every_ok) {
            const bool stale = box[0] != box_built[0] || box[1] != box_built[1] || box[2] != box_built[2];
            nflag = !p->check || stale ||
                    orc_needs_rebuild(n, x, y, z, xs.data(), ys.data(), zs.data(), box, p->skin, nullptr);
        }
        if (nflag) {
            rebuild();
            st->ago = 0;
            st->nrebuilds++;
        }
        // pre_force + force (the stress is always requested: requires_system_stress)
        std::memset(fx, 0, n * sizeof(double));
        std::memset(fy, 0, n * sizeof(double));
        std::memset(fz, 0, n * sizeof(double));
        double e = 0.0, vir[6] = {0, 0, 0, 0, 0, 0};
        pair_lj(n, x, y, z, box, hl.numneigh.data(), hl.firstneigh.data(), hl.neighbors.data(), p->eps,
                p->sigma, p->cutoff, 0, nullptr, nullptr, nullptr, nthreads, fx, fy, fz, &e, vir);
        for (int a = 0; a < 3; ++a)
            ns->vir[a] = vir[a];
        // final_integrate: P7
        m.kick(fx, fy, fz, half_dt);
        // end_of_step: P8-P11
        ns->ke_global = m.global_ke();
        m.scale_v_by_baro(half_dt);
        m.update_eps_dot(box, half_dt);
        m.nhc_atoms_half(half_dt);
        m.nhc_baro_half(half_dt);
        if (pe)
            pe[s] = e;
        if (ke)
            ke[s] = ns->ke_global;
        if (H)
            H[s] = m.hamiltonian(ns->ke_global, e, box);
        if (V)
            V[s] = box[0] * box[1] * box[2];
        st->step++;
    }
    return 0;
}

} // extern "C"
