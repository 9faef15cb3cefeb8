// atv_pme.cpp — CPU FP64 restatement of the reference's long-range
// electrostatics: smooth PME (B-splines, moduli, spreading, influence,
// gather, self and net-charge terms), the tabulated erfc real-space kernel
// of PairLJCutCoulLong and the closed-form CoulLong kernel.
//
// TEST INFRASTRUCTURE ONLY (the checker, never the product). The reference
// transforms with FFTW (c2c, FFTW_FORWARD = exp(-i...), unnormalised); FFTW is
// not in this image, so the 3-D transform here is this file's own mixed-radix
// Cooley-Tukey DFT (radices 2, 3, 5, 7, direct DFT for any other factor) with
// twiddles from cos/sin of the exact angle. It equals FFTW to rounding, not
// bitwise, which is the tolerance the reference's own PME tests use.
//
// Citations are to the reference tree (ELOISE-Wang0214/Atomiverse).

#include "atv_oracle.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cplx = std::complex<double>;

// include/atomiverse/core/units.h:134,173
constexpr double kPi = 3.14159265358979323846;
constexpr double kCoulomb = 14.399645;
// include/atomiverse/cpu/potentials/erfc_table.h:24 (2/sqrt(pi))
constexpr double kEwaldF = 1.1283791670955126;

// pme_spline.h:149-160 (pme_wrap)
inline int wrap_index(int g, int K) {
    if (g >= K) g -= K;
    if (g < 0) g += K;
    if (g >= K) g -= K;
    if (g < 0) g += K;
    return g;
}

// One-dimensional unnormalised DFT, X[k] = sum_j a[j] exp(sign 2 pi i j k / n),
// by recursive decimation in time over the smallest factor in {2,3,5,7}.
void fft1(std::vector<cplx>& a, int sign) {
    const int n = static_cast<int>(a.size());
    if (n <= 1)
        return;
    int p = n;
    for (int f : {2, 3, 5, 7})
        if (n % f == 0) {
            p = f;
            break;
        }
    const double base = sign * 2.0 * kPi / n;
    if (p == n) { // prime (or no small factor): direct O(n^2) sum
        std::vector<cplx> out(n);
        for (int k = 0; k < n; ++k) {
            cplx s = 0;
            for (int j = 0; j < n; ++j) {
                const long long jk = (static_cast<long long>(j) * k) % n;
                s += a[j] * std::polar(1.0, base * static_cast<double>(jk));
            }
            out[k] = s;
        }
        a.swap(out);
        return;
    }
    const int m = n / p;
    std::vector<std::vector<cplx>> sub(p, std::vector<cplx>(m));
    for (int r = 0; r < p; ++r)
        for (int j = 0; j < m; ++j)
            sub[r][j] = a[j * p + r];
    for (int r = 0; r < p; ++r)
        fft1(sub[r], sign);
    for (int k = 0; k < m; ++k)
        for (int q = 0; q < p; ++q) {
            const int kk = k + q * m;
            cplx s = 0;
            for (int r = 0; r < p; ++r) {
                const long long e = (static_cast<long long>(r) * kk) % n;
                s += sub[r][k] * std::polar(1.0, base * static_cast<double>(e));
            }
            a[kk] = s;
        }
}

// Three-dimensional transform of a grid laid out idx = k2 + K2 (k1 + K1 k0)
// (the reference's layout, pme_spline.h:165), axis by axis.
void fft3(cplx* g, const int K[3], int sign) {
    const int K0 = K[0], K1 = K[1], K2 = K[2];
    // axis 2 (contiguous)
#pragma omp parallel
    {
        std::vector<cplx> line;
#pragma omp for schedule(static)
        for (int r = 0; r < K0 * K1; ++r) {
            line.assign(g + static_cast<size_t>(r) * K2, g + static_cast<size_t>(r + 1) * K2);
            fft1(line, sign);
            std::memcpy(static_cast<void*>(g + static_cast<size_t>(r) * K2), line.data(),
                        K2 * sizeof(cplx));
        }
        line.resize(K1);
#pragma omp for schedule(static)
        for (int r = 0; r < K0 * K2; ++r) {
            const int k0 = r / K2, k2 = r % K2;
            for (int k1 = 0; k1 < K1; ++k1)
                line[k1] = g[k2 + K2 * (k1 + static_cast<size_t>(K1) * k0)];
            fft1(line, sign);
            for (int k1 = 0; k1 < K1; ++k1)
                g[k2 + K2 * (k1 + static_cast<size_t>(K1) * k0)] = line[k1];
        }
        line.resize(K0);
#pragma omp for schedule(static)
        for (int r = 0; r < K1 * K2; ++r) {
            const int k1 = r / K2, k2 = r % K2;
            for (int k0 = 0; k0 < K0; ++k0)
                line[k0] = g[k2 + K2 * (k1 + static_cast<size_t>(K1) * k0)];
            fft1(line, sign);
            for (int k0 = 0; k0 < K0; ++k0)
                g[k2 + K2 * (k1 + static_cast<size_t>(K1) * k0)] = line[k0];
        }
    }
}

} // namespace

extern "C" {

// good_fft_size (include/atomiverse/cpu/longrange/pme_spline.h:112-126)
int orc_good_fft_size(int n) {
    auto ok = [](int m) {
        for (int p : {2, 3, 5, 7})
            while (m % p == 0)
                m /= p;
        return m == 1;
    };
    if (n < 4)
        n = 4;
    while (!ok(n))
        ++n;
    return n;
}

// bspline, Essmann fill_bspline (pme_spline.h:29-55): theta[k] and dtheta[k]
// for k = 0..n-1 at fractional offset frac; order n-1 up-recursion, the
// derivative from the order n-1 values, then the last up-recursion.
void orc_bspline(double frac, int n, double* theta, double* dtheta) {
    theta[n - 1] = 0.0;
    theta[1] = frac;
    theta[0] = 1.0 - frac;
    for (int k = 3; k < n; ++k) {
        const double div = 1.0 / (k - 1);
        theta[k - 1] = div * frac * theta[k - 2];
        for (int j = 1; j < k - 1; ++j)
            theta[k - 1 - j] = div * ((frac + j) * theta[k - 2 - j] + (k - j - frac) * theta[k - 1 - j]);
        theta[0] = div * (1.0 - frac) * theta[0];
    }
    dtheta[0] = -theta[0];
    for (int j = 1; j < n; ++j)
        dtheta[j] = theta[j - 1] - theta[j];
    const double div = 1.0 / (n - 1);
    theta[n - 1] = div * frac * theta[n - 2];
    for (int j = 1; j < n - 1; ++j)
        theta[n - 1 - j] = div * ((frac + j) * theta[n - 2 - j] + (n - j - frac) * theta[n - 1 - j]);
    theta[0] = div * (1.0 - frac) * theta[0];
}

// spline_moduli (pme_spline.h:63-99): 1/|sum_k M_n(k+1) e^{2 pi i m k/K}|^2,
// zeros (even order at K/2) repaired by the mean of the two neighbours.
void orc_spline_moduli(int K, int n, double* mod) {
    std::vector<double> data(n, 0.0), dd(n, 0.0);
    orc_bspline(0.0, n, data.data(), dd.data());
    for (int i = 0; i < K; ++i) {
        double sr = 0, si = 0;
        for (int k = 0; k < n - 1; ++k) {
            const double arg = 2.0 * kPi * i * k / K;
            sr += data[k] * std::cos(arg);
            si += data[k] * std::sin(arg);
        }
        const double denom = sr * sr + si * si;
        mod[i] = denom > 1e-12 ? 1.0 / denom : 0.0;
    }
    for (int i = 0; i < K; ++i)
        if (mod[i] == 0.0)
            mod[i] = 0.5 * (mod[(i - 1 + K) % K] + mod[(i + 1) % K]);
}

// EwaldSum::recommend_alpha (include/atomiverse/cpu/longrange/ewald_sum.h:58-60)
double orc_ewald_alpha(double cutoff, double accuracy) {
    return (1.35 - 0.15 * std::log(accuracy)) / cutoff;
}

// KSpacePME::ensure_grid mesh size (src/cpu/longrange/pme.cpp:66-74)
void orc_pme_grid(const double box[3], double spacing, int order, int K[3]) {
    for (int d = 0; d < 3; ++d) {
        K[d] = orc_good_fft_size(static_cast<int>(std::ceil(box[d] / spacing)));
        if (K[d] < order + 1)
            K[d] = orc_good_fft_size(order + 1);
    }
}

void orc_fft3d(int K0, int K1, int K2, double* data, int sign) {
    const int K[3] = {K0, K1, K2};
    fft3(reinterpret_cast<cplx*>(data), K, sign);
}

// KSpacePME::compute (src/cpu/longrange/pme.cpp:147-253) for an orthorhombic
// box on one rank:
//   pme_fill_atom_splines (pme_spline.h:131-146) with inv_cell = diag(1/L),
//   pme_spread_charges (pme_spline.h:170-206) onto a zeroed complex grid,
//   forward c2c transform, pme_build_influence_tables +
//   pme_apply_influence_tables (pme_influence.h:115-222), backward c2c,
//   pme_interpolate_forces (pme_influence.h:226-281), self and net-charge
//   energies (pme.cpp:222-246).
// Forces are ACCUMULATED into f*, energy into *energy (self first, then
// recip + net, the reference's order), the Voigt virial into virial[6].
// parts[3] (optional) receives e_recip, e_self, e_net.
int orc_pme(int64_t n, const double* x, const double* y, const double* z, const double* q,
            const double box[3], double alpha, const int K[3], int order, double* fx, double* fy,
            double* fz, double* energy, double* virial, double* parts) {
    if (order < 3 || K[0] < order + 1 || K[1] < order + 1 || K[2] < order + 1 || alpha <= 0.0)
        return -1;
    const int K0 = K[0], K1 = K[1], K2 = K[2];
    const int64_t ntot = static_cast<int64_t>(K0) * K1 * K2;
    // Box::recompute_cache_ (src/cpu/frame/box.cpp:68-81): inv_cell = diag(1/L)
    const double ic[3] = {1.0 / box[0], 1.0 / box[1], 1.0 / box[2]};
    const double vol = box[0] * box[1] * box[2];
    const int nrd = order;
    const int nm1 = nrd - 1;

    std::vector<double> mod[3];
    for (int d = 0; d < 3; ++d) {
        mod[d].resize(K[d]);
        orc_spline_moduli(K[d], nrd, mod[d].data());
    }

    // per-atom splines
    std::vector<int> base(3 * n);
    std::vector<double> theta(3 * nrd * n), dtheta(3 * nrd * n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const double r[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; ++d) {
            // inv_cell(0,d)*x + inv_cell(1,d)*y + inv_cell(2,d)*z with the
            // off-diagonal zeros: the products by zero add exactly nothing.
            const double s = ic[d] * r[d];
            double u = s * K[d];
            u -= std::floor(u / K[d]) * K[d];
            const int u0 = static_cast<int>(std::floor(u));
            const double frac = u - u0;
            orc_bspline(frac, nrd, &theta[(3 * i + d) * nrd], &dtheta[(3 * i + d) * nrd]);
            base[3 * i + d] = u0;
        }
    }

    // spread (serial, ascending atom order, as the reference)
    std::vector<cplx> grid(ntot, cplx(0, 0));
    for (int64_t i = 0; i < n; ++i) {
        const double qi = q[i];
        if (qi == 0.0)
            continue;
        const double* thx = &theta[(3 * i + 0) * nrd];
        const double* thy = &theta[(3 * i + 1) * nrd];
        const double* thz = &theta[(3 * i + 2) * nrd];
        const int b0 = base[3 * i + 0] - nm1, b1 = base[3 * i + 1] - nm1, b2 = base[3 * i + 2] - nm1;
        for (int a = 0; a < nrd; ++a) {
            const int g0 = wrap_index(b0 + a, K0);
            const double wa = qi * thx[a];
            for (int b = 0; b < nrd; ++b) {
                const int g1 = wrap_index(b1 + b, K1);
                const double wab = wa * thy[b];
                const int64_t plane = static_cast<int64_t>(K2) * (g1 + static_cast<int64_t>(K1) * g0);
                for (int c = 0; c < nrd; ++c) {
                    const int g2 = wrap_index(b2 + c, K2);
                    grid[plane + g2].real(grid[plane + g2].real() + wab * thz[c]);
                }
            }
        }
    }

    fft3(grid.data(), K, -1);

    // influence tables + application, fused per mode in table order
    const double inv_pi2_a2 = kPi * kPi / (alpha * alpha);
    double e_recip = 0, s_xx = 0, s_yy = 0, s_zz = 0, s_xy = 0, s_xz = 0, s_yz = 0;
    for (int m0 = 0; m0 < K0; ++m0) {
        const int h0 = (m0 <= K0 / 2) ? m0 : m0 - K0;
        const double bm0 = mod[0][m0];
        for (int m1 = 0; m1 < K1; ++m1) {
            const int h1 = (m1 <= K1 / 2) ? m1 : m1 - K1;
            const double bm01 = bm0 * mod[1][m1];
            for (int m2 = 0; m2 < K2; ++m2) {
                const int64_t idx = m2 + static_cast<int64_t>(K2) * (m1 + static_cast<int64_t>(K1) * m0);
                if (m0 == 0 && m1 == 0 && m2 == 0) {
                    grid[idx] = 0;
                    continue;
                }
                const int h2 = (m2 <= K2 / 2) ? m2 : m2 - K2;
                // m = sum_d h_d (column d of inv_cell); diagonal inv_cell
                const double mx = h0 * ic[0];
                const double my = h1 * ic[1];
                const double mz = h2 * ic[2];
                const double msq = mx * mx + my * my + mz * mz;
                const double inv_msq = 1.0 / msq;
                const double B = bm01 * mod[2][m2];
                const double H = std::exp(-inv_pi2_a2 * msq) * inv_msq * B;
                if (H == 0.0) {
                    grid[idx] = 0;
                    continue;
                }
                const cplx Qm = grid[idx];
                const double q2 = Qm.real() * Qm.real() + Qm.imag() * Qm.imag();
                e_recip += H * q2;
                const double vt = -2.0 * (inv_msq + inv_pi2_a2);
                const double ug = H * q2;
                s_xx += ug * (1.0 + vt * mx * mx);
                s_yy += ug * (1.0 + vt * my * my);
                s_zz += ug * (1.0 + vt * mz * mz);
                s_xy += ug * (vt * mx * my);
                s_xz += ug * (vt * mx * mz);
                s_yz += ug * (vt * my * mz);
                grid[idx] = H * Qm;
            }
        }
    }
    const double escale = kCoulomb / (2.0 * kPi * vol);
    e_recip *= escale;

    fft3(grid.data(), K, +1);

    // gather
    const double fscale = kCoulomb / (kPi * vol);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const double qi = q[i];
        if (qi == 0.0)
            continue;
        const double* thx = &theta[(3 * i + 0) * nrd];
        const double* thy = &theta[(3 * i + 1) * nrd];
        const double* thz = &theta[(3 * i + 2) * nrd];
        const double* dthx = &dtheta[(3 * i + 0) * nrd];
        const double* dthy = &dtheta[(3 * i + 1) * nrd];
        const double* dthz = &dtheta[(3 * i + 2) * nrd];
        const int b0 = base[3 * i + 0] - nm1, b1 = base[3 * i + 1] - nm1, b2 = base[3 * i + 2] - nm1;
        double g_d[3] = {0, 0, 0};
        for (int a = 0; a < nrd; ++a) {
            const int g0 = wrap_index(b0 + a, K0);
            for (int b = 0; b < nrd; ++b) {
                const int g1 = wrap_index(b1 + b, K1);
                const int64_t plane = static_cast<int64_t>(K2) * (g1 + static_cast<int64_t>(K1) * g0);
                for (int c = 0; c < nrd; ++c) {
                    const int g2 = wrap_index(b2 + c, K2);
                    const double cval = grid[plane + g2].real();
                    g_d[0] += cval * dthx[a] * thy[b] * thz[c];
                    g_d[1] += cval * thx[a] * dthy[b] * thz[c];
                    g_d[2] += cval * thx[a] * thy[b] * dthz[c];
                }
            }
        }
        // sum_d g_d K_d inv_cell(a, d): only d == a survives for a diagonal
        // inv_cell (the zero products add exactly nothing).
        const double fxi = g_d[0] * K[0] * ic[0];
        const double fyi = g_d[1] * K[1] * ic[1];
        const double fzi = g_d[2] * K[2] * ic[2];
        fx[i] += -fscale * qi * fxi;
        fy[i] += -fscale * qi * fyi;
        fz[i] += -fscale * qi * fzi;
    }

    double qsq = 0, qsum = 0;
    for (int64_t i = 0; i < n; ++i) {
        qsq += q[i] * q[i];
        qsum += q[i];
    }
    const double e_self = -kCoulomb * alpha / std::sqrt(kPi) * qsq;
    const double e_net = -kCoulomb * kPi / (2.0 * vol * alpha * alpha) * qsum * qsum;
    if (energy) {
        *energy += e_self;
        *energy += e_recip + e_net;
    }
    if (virial) {
        virial[0] += s_xx * escale;
        virial[1] += s_yy * escale;
        virial[2] += s_zz * escale;
        virial[3] += s_xy * escale;
        virial[4] += s_xz * escale;
        virial[5] += s_yz * escale;
    }
    if (parts) {
        parts[0] = e_recip;
        parts[1] = e_self;
        parts[2] = e_net;
    }
    return 0;
}

// ErfcTable::build (include/atomiverse/cpu/potentials/erfc_table.h:44-72):
// nbin+1 entries over r^2 in [0, cutoff^2], entry 0 copies entry 1.
void orc_erfc_table(double alpha, double cutoff, int nbin, double* e_tab, double* f_tab) {
    const double cutoff_sq = cutoff * cutoff;
    const double dr2 = cutoff_sq / static_cast<double>(nbin);
    for (int i = 1; i <= nbin; ++i) {
        const double r2 = i * dr2;
        const double r = std::sqrt(r2);
        const double grij = alpha * r;
        const double expm2 = std::exp(-grij * grij);
        const double erfcv = std::erfc(grij);
        e_tab[i] = erfcv / r;
        f_tab[i] = (erfcv + kEwaldF * grij * expm2) / (r2 * r);
    }
    e_tab[0] = e_tab[1];
    f_tab[0] = f_tab[1];
}

// PairLJCutCoulLong::finalize_type_table_ (src/cpu/potentials/
// pair_lj_cut_coul_long.cpp:132-186): Lorentz-Berthelot for unset
// off-diagonals whose diagonals are set, (eps 0, sigma 1) otherwise;
// lj1..lj4 = 48 eps s^12, 24 eps s^6, 4 eps s^12, 4 eps s^6; energy shift at
// the cutoff when eps > 0. eps/sigma are updated in place for mixed entries.
void orc_ljcoul_tables(int ntypes, double cutoff, double* eps, double* sigma, const int* set,
                       double* lj1, double* lj2, double* lj3, double* lj4, double* shift) {
    for (int a = 0; a < ntypes; ++a)
        for (int b = 0; b < ntypes; ++b) {
            const int ab = a * ntypes + b;
            double e = eps[ab], s = sigma[ab];
            if (!set[ab]) {
                const int aa = a * ntypes + a, bb = b * ntypes + b;
                if (set[aa] && set[bb]) {
                    e = std::sqrt(eps[aa] * eps[bb]);
                    s = 0.5 * (sigma[aa] + sigma[bb]);
                } else {
                    e = 0.0;
                    s = 1.0;
                }
                eps[ab] = e;
                sigma[ab] = s;
            }
            const double s2 = s * s;
            const double s6 = s2 * s2 * s2;
            const double s12 = s6 * s6;
            lj1[ab] = 48.0 * e * s12;
            lj2[ab] = 24.0 * e * s6;
            lj3[ab] = 4.0 * e * s12;
            lj4[ab] = 4.0 * e * s6;
            if (e > 0.0 && cutoff > 0.0) {
                const double rc2 = cutoff * cutoff;
                const double r6 = rc2 * rc2 * rc2;
                const double r6inv = 1.0 / r6;
                shift[ab] = lj3[ab] * r6inv * r6inv - lj4[ab] * r6inv;
            } else {
                shift[ab] = 0.0;
            }
        }
}

// PairLJCutCoulLong::compute (src/cpu/potentials/pair_lj_cut_coul_long.cpp:
// 276-571) on a half CSR without shifts: min image by nearbyint(d / L) with
// 1/L hoisted, r2 = dx^2 + dy^2 + dz^2, survivors r2 < rc^2 and r2 != 0,
// LJ from lj1..lj4 and the shift, Coulomb from the erfc table interpolated
// linearly in r^2, special scales 1 (no topology), pairs with zero force and
// energy skipped. Tail correction off (the reference default). Forces are
// accumulated; energy and virial are added to *energy and virial[6].
// Tables: lj*/shift from orc_ljcoul_tables, e_tab/f_tab from orc_erfc_table.
void orc_pair_lj_coul_long(int64_t n, const double* x, const double* y, const double* z,
                           const double* q, const int* type, const double box[3],
                           const int* numneigh, const int64_t* firstneigh, const int* neighbors,
                           int ntypes, const double* lj1, const double* lj2, const double* lj3,
                           const double* lj4, const double* shift, double cutoff,
                           const double* e_tab, const double* f_tab, int nbin, int nthreads,
                           double* fx, double* fy, double* fz, double* energy, double* virial) {
    const double rc_sq = cutoff * cutoff;
    const double dr2 = rc_sq / static_cast<double>(nbin);
    const double inv_dr2 = 1.0 / dr2;
    const double lx = box[0], ly = box[1], lz = box[2];
    const double inv_lx = 1.0 / lx, inv_ly = 1.0 / ly, inv_lz = 1.0 / lz;
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
            const double xi = x[i], yi = y[i], zi = z[i];
            const double qi_scale = kCoulomb * q[i];
            const int ti = type[i];
            const int jnum = numneigh[i];
            const int64_t off = firstneigh[i];
            double fi_x = 0, fi_y = 0, fi_z = 0;
            for (int jj = 0; jj < jnum; ++jj) {
                const int j = neighbors[off + jj];
                double dx = xi - x[j];
                double dy = yi - y[j];
                double dz = zi - z[j];
                dx -= lx * std::nearbyint(dx * inv_lx);
                dy -= ly * std::nearbyint(dy * inv_ly);
                dz -= lz * std::nearbyint(dz * inv_lz);
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 >= rc_sq || r2 == 0.0)
                    continue;
                double fpair = 0.0, epair = 0.0;
                const int ab = ti * ntypes + type[j];
                const double a1 = lj1[ab];
                if (a1 != 0.0) {
                    const double r2inv = 1.0 / r2;
                    const double r6inv = r2inv * r2inv * r2inv;
                    fpair += 1.0 * r2inv * (a1 * r6inv * r6inv - lj2[ab] * r6inv);
                    epair += 1.0 * (lj3[ab] * r6inv * r6inv - lj4[ab] * r6inv - shift[ab]);
                }
                if (qi_scale != 0.0) {
                    const double qj = q[j];
                    if (qj != 0.0) {
                        const double xb = r2 * inv_dr2;
                        int bi = static_cast<int>(xb);
                        if (bi >= nbin)
                            bi = nbin - 1;
                        const double t = xb - static_cast<double>(bi);
                        const double et = e_tab[bi] + t * (e_tab[bi + 1] - e_tab[bi]);
                        const double ft = f_tab[bi] + t * (f_tab[bi + 1] - f_tab[bi]);
                        const double qq = qi_scale * qj;
                        fpair += qq * ft;
                        epair += qq * et;
                    }
                }
                if (fpair == 0.0 && epair == 0.0)
                    continue;
                const double fxp = dx * fpair, fyp = dy * fpair, fzp = dz * fpair;
                fi_x += fxp;
                fi_y += fyp;
                fi_z += fzp;
                fxt[j] -= fxp;
                fyt[j] -= fyp;
                fzt[j] -= fzp;
                e_tot += epair;
                v_xx += dx * fxp;
                v_yy += dy * fyp;
                v_zz += dz * fzp;
                v_xy += dx * fyp;
                v_xz += dx * fzp;
                v_yz += dy * fzp;
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
    if (energy)
        *energy += e_tot;
    if (virial) {
        virial[0] += v_xx;
        virial[1] += v_yy;
        virial[2] += v_zz;
        virial[3] += v_xy;
        virial[4] += v_xz;
        virial[5] += v_yz;
    }
}

// CoulLong::compute (src/cpu/potentials/coul_long.cpp:120-236): closed-form
// erfc real-space Coulomb on a half CSR, min image by nearbyint, no specials.
void orc_coul_long(int64_t n, const double* x, const double* y, const double* z, const double* q,
                   const double box[3], const int* numneigh, const int64_t* firstneigh,
                   const int* neighbors, double alpha, double cutoff, double* fx, double* fy,
                   double* fz, double* energy, double* virial) {
    const double rc_sq = cutoff * cutoff;
    const double lx = box[0], ly = box[1], lz = box[2];
    const double inv_lx = 1.0 / lx, inv_ly = 1.0 / ly, inv_lz = 1.0 / lz;
    double e_tot = 0, v[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        const double xi = x[i], yi = y[i], zi = z[i];
        const double qi = q[i];
        if (qi == 0.0)
            continue;
        double fi_x = 0, fi_y = 0, fi_z = 0;
        for (int jj = 0; jj < numneigh[i]; ++jj) {
            const int j = neighbors[firstneigh[i] + jj];
            double dx = xi - x[j], dy = yi - y[j], dz = zi - z[j];
            dx -= lx * std::nearbyint(dx * inv_lx);
            dy -= ly * std::nearbyint(dy * inv_ly);
            dz -= lz * std::nearbyint(dz * inv_lz);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 >= rc_sq || r2 == 0.0)
                continue;
            const double qj = q[j];
            if (qj == 0.0)
                continue;
            const double r = std::sqrt(r2);
            const double r2inv = 1.0 / r2;
            const double grij = alpha * r;
            const double expm2 = std::exp(-grij * grij);
            const double erfcv = std::erfc(grij);
            const double prefactor = kCoulomb * qi * qj / r;
            const double forcecoul = prefactor * (erfcv + kEwaldF * grij * expm2);
            const double fpair = forcecoul * r2inv;
            const double fxp = dx * fpair, fyp = dy * fpair, fzp = dz * fpair;
            fi_x += fxp;
            fi_y += fyp;
            fi_z += fzp;
            fx[j] -= fxp;
            fy[j] -= fyp;
            fz[j] -= fzp;
            v[0] += dx * fxp;
            v[1] += dy * fyp;
            v[2] += dz * fzp;
            v[3] += dx * fyp;
            v[4] += dx * fzp;
            v[5] += dy * fzp;
            e_tot += prefactor * erfcv;
        }
        fx[i] += fi_x;
        fy[i] += fi_y;
        fz[i] += fi_z;
    }
    if (energy)
        *energy += e_tot;
    if (virial)
        for (int k = 0; k < 6; ++k)
            virial[k] += v[k];
}

} // extern "C"
