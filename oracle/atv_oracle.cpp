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
