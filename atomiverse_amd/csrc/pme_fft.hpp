// pme_fft.hpp — the FP32 PME mesh transforms for power-of-two meshes, written
// for gfx950 in place of the FFT library's 3D R2C / C2R (which other mesh
// sizes and FP64 keep). The reference transforms with FFTW r2c / c2r
// (src/cpu/longrange/pme.cpp:152-253); the same unnormalised transforms are
// computed here as five passes over the half spectrum [K0][K1][K2/2+1], each
// reading and writing the mesh once:
//   A  z-lines: fixed-point mesh -> real (the unfix, which also clears the
//      fixed-point mesh for the next spread), real -> half spectrum (an
//      M = K2/2 point complex FFT of the even/odd pairs, then the split into
//      the K2/2+1 modes);
//   B  y-lines, forward;
//   C  x-lines: forward, influence function (energy and virial sums in a
//      fixed order, as the separate influence kernel does), inverse — the
//      line never leaves LDS in between;
//   D  y-lines, inverse;
//   E  z-lines: half spectrum -> real (the inverse split, an M point inverse
//      FFT, the even/odd unpacking) into the real mesh the gather reads.
// Lines are transformed in LDS by radix-2 butterflies: decimation in time
// from bit-reversed order to natural order, or decimation in frequency from
// natural to bit-reversed order, so each pass permutes once, in its load or
// store addresses. Every line is owned by the lanes of one wavefront, so the
// stages need no workgroup barrier (LDS operations of a wavefront complete
// in issue order). Passes B, C and D move tiles of kFftTile columns, one
// wavefront per line (see TileMap). Twiddles come from host-built tables
// (long double sin/cos, rounded once).

#pragma once

#include "atvh_internal.hpp"

#include <type_traits>

namespace atvh {
namespace pmefft {

constexpr int kFftThreads = 256;  // four wavefronts per workgroup
constexpr int kFftTile = 16;      // columns per tile of passes B, C, D
constexpr int kFftMaxLog2 = 8;    // lines of at most 256 points
constexpr int kFftLineGrid = 2048; // workgroups of passes A and E (grid-stride)
constexpr int kFftTileGrid = 512;  // workgroups of passes B, C and D (grid-stride, two per CU)

template <class T> struct Cplx;
template <> struct Cplx<float> { using type = float2; };
template <> struct Cplx<double> { using type = double2; };

template <class T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b) {
    T2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <class T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) {
    T2 r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}
template <class T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) {
    T2 r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    return r;
}
template <class T2> __device__ __forceinline__ T2 cconj(T2 a) {
    a.y = -a.y;
    return a;
}

__device__ __forceinline__ int bitrev(int i, int log2n) {
    return static_cast<int>(__builtin_bitreverse32(static_cast<unsigned>(i)) >> (32 - log2n));
}

// Decimation in time over one N-point line in LDS: bit-reversed in, natural
// out. tw[k * ts] = exp(-2 pi i k / N) for k < N/2 (ts: stride into a table
// of a longer line); INV uses the conjugates. Lanes t of nt (one wavefront).
template <class T2, int LOG2N, bool INV>
__device__ __forceinline__ void dit(T2* s, const T2* tw, int ts, int t, int nt) {
    constexpr int N = 1 << LOG2N;
#pragma unroll
    for (int st = 0; st < LOG2N; ++st) {
        const int half = 1 << st;
        for (int b = t; b < N / 2; b += nt) {
            const int j = b & (half - 1);
            const int i0 = ((b >> st) << (st + 1)) + j, i1 = i0 + half;
            T2 w = tw[(j << (LOG2N - 1 - st)) * ts];
            if (INV)
                w.y = -w.y;
            const T2 u = s[i0], v = cmul(w, s[i1]);
            s[i0] = cadd(u, v);
            s[i1] = csub(u, v);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Decimation in frequency: natural in, bit-reversed out.
template <class T2, int LOG2N, bool INV>
__device__ __forceinline__ void dif(T2* s, const T2* tw, int ts, int t, int nt) {
    constexpr int N = 1 << LOG2N;
#pragma unroll
    for (int st = LOG2N - 1; st >= 0; --st) {
        const int half = 1 << st;
        for (int b = t; b < N / 2; b += nt) {
            const int j = b & (half - 1);
            const int i0 = ((b >> st) << (st + 1)) + j, i1 = i0 + half;
            T2 w = tw[(j << (LOG2N - 1 - st)) * ts];
            if (INV)
                w.y = -w.y;
            const T2 u = s[i0], v = s[i1];
            s[i0] = cadd(u, v);
            s[i1] = cmul(w, csub(u, v));
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Rows of the half spectrum are padded from Kh = K2/2+1 to ks = a multiple
// of kFftTile complex values, so every tile row is whole aligned cache lines.
__host__ __device__ constexpr int fft_row(int Kh) { return (Kh + kFftTile - 1) / kFftTile * kFftTile; }

struct FftArgs {
    int K0, K1, K2, Kh;
    int ks; // row stride of the half spectrum (fft_row(Kh))
    double inv_fix;
    unsigned long long* fix; // fixed-point mesh [K0][K1][K2] (pass A reads and clears it)
    void* real;              // real mesh [K0][K1][K2] (pass E writes it)
    void* cplx;              // half spectrum [K0][K1][ks] (columns Kh..ks-1 unused)
    const void* tw0;         // exp(-2 pi i k / K0), k <= K0/2
    const void* tw1;         // exp(-2 pi i k / K1), k <= K1/2
    const void* tw2;         // exp(-2 pi i k / K2), k <= K2/2
    // pass C: influence function and its energy / virial sums
    const double* H;
    double ic0, ic1, ic2, pi2_a2;
    double* partials;
};

// Pass A: LOG2M = log2(K2/2). One line per nt = min(64, M/2) lanes.
template <class T, int LOG2M>
__global__ __launch_bounds__(kFftThreads) void k_fft_r2c_z(FftArgs a) {
    using T2 = typename Cplx<T>::type;
    constexpr int M = 1 << LOG2M;
    constexpr int NT = M / 2 < 64 ? M / 2 : 64;
    constexpr int LPW = 64 / NT; // lines per wavefront
    __shared__ T2 buf[kFftThreads / 64][LPW][M];
    __shared__ T2 tw[M + 1];
    const T2* gtw = static_cast<const T2*>(a.tw2);
    for (int k = threadIdx.x; k <= M; k += blockDim.x)
        tw[k] = gtw[k];
    __syncthreads();
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int sub = lane / NT, t = lane % NT;
    T2* s = buf[wid][sub];
    const int nlines = a.K0 * a.K1;
    const int wstride = gridDim.x * (kFftThreads / 64) * LPW;
    T2* out = static_cast<T2*>(a.cplx);
    for (int l0 = (blockIdx.x * (kFftThreads / 64) + wid) * LPW; l0 < nlines; l0 += wstride) {
        const int line = l0 + sub; // nlines is a multiple of LPW
        ulonglong2* fx = reinterpret_cast<ulonglong2*>(a.fix + (int64_t)line * (2 * M));
        for (int m = t; m < M; m += NT) {
            const ulonglong2 v = fx[m];
            T2 z;
            z.x = static_cast<T>(static_cast<double>(static_cast<long long>(v.x)) * a.inv_fix);
            z.y = static_cast<T>(static_cast<double>(static_cast<long long>(v.y)) * a.inv_fix);
            fx[m] = make_ulonglong2(0ull, 0ull);
            s[bitrev(m, LOG2M)] = z;
        }
        __builtin_amdgcn_wave_barrier();
        dit<T2, LOG2M, false>(s, tw, 2, t, NT);
        // X[k] = (Z[k] + conj Z[M-k]) / 2 - i W^k (Z[k] - conj Z[M-k]) / 2, k = 0..M
        T2* o = out + (int64_t)line * a.ks;
        for (int k = t; k <= M; k += NT) {
            const T2 zk = s[k & (M - 1)], zc = cconj(s[(M - k) & (M - 1)]);
            const T2 e = cadd(zk, zc), d = csub(zk, zc);
            const T2 w = tw[k];
            // -i w d = (w.x d.y + w.y d.x) - i (w.x d.x - w.y d.y)
            const T2 wd = cmul(w, d);
            T2 x;
            x.x = T(0.5) * (e.x + wd.y);
            x.y = T(0.5) * (e.y - wd.x);
            o[k] = x;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Pass E: the inverse of pass A (unnormalised: K2 times the inverse).
template <class T, int LOG2M>
__global__ __launch_bounds__(kFftThreads) void k_fft_c2r_z(FftArgs a) {
    using T2 = typename Cplx<T>::type;
    constexpr int M = 1 << LOG2M;
    constexpr int NT = M / 2 < 64 ? M / 2 : 64;
    constexpr int LPW = 64 / NT;
    __shared__ T2 buf[kFftThreads / 64][LPW][M];
    __shared__ T2 tw[M + 1];
    const T2* gtw = static_cast<const T2*>(a.tw2);
    for (int k = threadIdx.x; k <= M; k += blockDim.x)
        tw[k] = gtw[k];
    __syncthreads();
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int sub = lane / NT, t = lane % NT;
    T2* s = buf[wid][sub];
    const int nlines = a.K0 * a.K1;
    const int wstride = gridDim.x * (kFftThreads / 64) * LPW;
    const T2* in = static_cast<const T2*>(a.cplx);
    T2* real = static_cast<T2*>(a.real);
    for (int l0 = (blockIdx.x * (kFftThreads / 64) + wid) * LPW; l0 < nlines; l0 += wstride) {
        const int line = l0 + sub;
        const T2* x = in + (int64_t)line * a.ks;
        // Z[k] = (X[k] + conj X[M-k]) + i conj(W^k) (X[k] - conj X[M-k]), k < M
        for (int k = t; k < M; k += NT) {
            const T2 xk = x[k], xc = cconj(x[M - k]);
            const T2 e = cadd(xk, xc), d = csub(xk, xc);
            const T2 wd = cmul(cconj(tw[k]), d);
            T2 z;
            z.x = e.x - wd.y;
            z.y = e.y + wd.x;
            s[k] = z;
        }
        __builtin_amdgcn_wave_barrier();
        dif<T2, LOG2M, true>(s, tw, 2, t, NT);
        T2* r = real + (int64_t)line * M; // the real line as M (even, odd) pairs
        for (int m = t; m < M; m += NT)
            r[m] = s[bitrev(m, LOG2M)];
        __builtin_amdgcn_wave_barrier();
    }
}

// Tiles of passes B, C and D. The half spectrum's columns m2 < K2/2 are cut
// into full tiles of kFftTile adjacent columns at one outer index (i0 for
// y-lines, m1 for x-lines): every row of such a tile is one contiguous
// segment. The last column, m2 = K2/2, is gathered kFftTile outer indices at
// a time into tiles of its own, so no workgroup transforms empty columns.
// Element j of column c of a tile sits at base + c * cs + j * ls.
struct TileMap {
    int64_t base, cs;
    int outer0, m2;  // outer index and m2 of column 0
    bool nyquist;    // columns step the outer index (else m2)
};

__device__ __forceinline__ TileMap tile_map(int tile, int nouter, int Kh, int64_t outer_stride) {
    const int full = (Kh - 1) / kFftTile; // full tiles per outer index
    TileMap t;
    if (tile < nouter * full) {
        t.outer0 = tile / full;
        t.m2 = (tile % full) * kFftTile;
        t.nyquist = false;
        t.base = (int64_t)t.outer0 * outer_stride + t.m2;
        t.cs = 1;
    } else {
        t.outer0 = (tile - nouter * full) * kFftTile;
        t.m2 = Kh - 1;
        t.nyquist = true;
        t.base = (int64_t)t.outer0 * outer_stride + t.m2;
        t.cs = outer_stride;
    }
    return t;
}

__device__ __forceinline__ int fft_tiles(int nouter, int Kh) {
    return nouter * ((Kh - 1) / kFftTile) + nouter / kFftTile;
}

constexpr int kFftTileThreads = 64 * kFftTile; // one wavefront per line of a tile
// FP32 tiles load the workgroup's next tile while the current one transforms
// (11.6 vs 14.8 us per y pass at 128^3).
template <class T> constexpr bool kFftPrefetch = std::is_same<T, float>::value;

// Passes B and D: y-lines (outer index i0, element stride ks). Each thread
// moves E = kFftTile N / kFftTileThreads elements of a tile: element e of
// the tile is (j = e / kFftTile, column e % kFftTile). The next tile's
// elements are loaded into registers while the current one transforms.
template <class T, int LOG2N, bool INV>
__global__ __launch_bounds__(kFftTileThreads) void k_fft_y(FftArgs a) {
    using T2 = typename Cplx<T>::type;
    constexpr int N = 1 << LOG2N;
    constexpr int NT = N / 2 < 64 ? N / 2 : 64;
    constexpr int LPW = 64 / NT;
    constexpr int E = (kFftTile * N + kFftTileThreads - 1) / kFftTileThreads;
    __shared__ T2 buf[kFftTile][N + 1]; // + 1: the tile's columns fall in different LDS banks
    __shared__ T2 tw[N / 2];
    const T2* gtw = static_cast<const T2*>(a.tw1);
    for (int k = threadIdx.x; k < N / 2; k += blockDim.x)
        tw[k] = gtw[k];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col_w = wid * LPW + lane / NT, t = lane % NT;
    const int64_t ls = a.ks, outer_stride = (int64_t)a.K1 * a.ks;
    const int ntiles = fft_tiles(a.K0, a.Kh);
    T2* c = static_cast<T2*>(a.cplx);
    T2 r[E];
    auto load = [&](int tile) {
        const TileMap m = tile_map(tile, a.K0, a.Kh, outer_stride);
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = threadIdx.x + k * kFftTileThreads;
            if (e < kFftTile * N)
                r[k] = c[m.base + (e % kFftTile) * m.cs + (e / kFftTile) * ls];
        }
    };
    if (kFftPrefetch<T> && blockIdx.x < ntiles)
        load(blockIdx.x);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const TileMap m = tile_map(tile, a.K0, a.Kh, outer_stride);
        __syncthreads(); // the previous tile's stores have read buf
        if (!kFftPrefetch<T>)
            load(tile);
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = threadIdx.x + k * kFftTileThreads;
            if (e < kFftTile * N)
                buf[e % kFftTile][bitrev(e / kFftTile, LOG2N)] = r[k];
        }
        if (kFftPrefetch<T> && tile + gridDim.x < ntiles)
            load(tile + gridDim.x);
        __syncthreads();
        if (col_w < kFftTile)
            dit<T2, LOG2N, INV>(buf[col_w], tw, 1, t, NT);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = threadIdx.x + k * kFftTileThreads;
            if (e < kFftTile * N)
                c[m.base + (e % kFftTile) * m.cs + (e / kFftTile) * ls] = buf[e % kFftTile][e / kFftTile];
        }
    }
}

// Pass C: x-lines (outer index m1, element stride K1 ks) — forward
// (decimation in frequency: the modes come out bit-reversed), influence
// function and its sums (pme_apply_influence_tables, pme_influence.h:179-222,
// folded over the Hermitian partners as the separate influence kernel does),
// inverse (decimation in time: bit-reversed in, natural out). Element e of a
// tile is (j = e / kFftTile, column e % kFftTile) throughout; the next tile
// and this tile's influence values are loaded while the current one works.
template <class T, int LOG2N, bool EV>
__global__ __launch_bounds__(kFftTileThreads) void k_fft_x_influence(FftArgs a) {
    using T2 = typename Cplx<T>::type;
    constexpr int N = 1 << LOG2N;
    constexpr int NT = N / 2 < 64 ? N / 2 : 64;
    constexpr int LPW = 64 / NT;
    constexpr int E = (kFftTile * N + kFftTileThreads - 1) / kFftTileThreads;
    __shared__ T2 buf[kFftTile][N + 1]; // + 1: the tile's columns fall in different LDS banks
    __shared__ T2 tw[N / 2];
    const T2* gtw = static_cast<const T2*>(a.tw0);
    for (int k = threadIdx.x; k < N / 2; k += blockDim.x)
        tw[k] = gtw[k];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col_w = wid * LPW + lane / NT, t = lane % NT;
    const int64_t ls = (int64_t)a.K1 * a.ks, outer_stride = a.ks;
    const int ntiles = fft_tiles(a.K1, a.Kh);
    T2* c = static_cast<T2*>(a.cplx);
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    T2 r[E];
    auto load = [&](int tile) {
        const TileMap m = tile_map(tile, a.K1, a.Kh, outer_stride);
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = threadIdx.x + k * kFftTileThreads;
            if (e < kFftTile * N)
                r[k] = c[m.base + (e % kFftTile) * m.cs + (e / kFftTile) * ls];
        }
    };
    if (kFftPrefetch<T> && blockIdx.x < ntiles)
        load(blockIdx.x);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const TileMap m = tile_map(tile, a.K1, a.Kh, outer_stride);
        __syncthreads();
        if (!kFftPrefetch<T>)
            load(tile);
        double hv[E];
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = threadIdx.x + k * kFftTileThreads;
            hv[k] = 0.0;
            if (e < kFftTile * N) {
                const int col = e % kFftTile, j = e / kFftTile;
                buf[col][j] = r[k];
                // this element's mode after the forward pass: m0 = bitrev(j)
                const int m1 = m.nyquist ? m.outer0 + col : m.outer0;
                const int m2 = m.nyquist ? m.m2 : m.m2 + col;
                hv[k] = a.H[((int64_t)bitrev(j, LOG2N) * a.K1 + m1) * a.Kh + m2];
            }
        }
        if (kFftPrefetch<T> && tile + gridDim.x < ntiles)
            load(tile + gridDim.x);
        __syncthreads();
        if (col_w < kFftTile)
            dif<T2, LOG2N, false>(buf[col_w], tw, 1, t, NT);
        __syncthreads();
        // buf[col][p] holds mode m0 = bitrev(p)
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = threadIdx.x + k * kFftTileThreads;
            if (e >= kFftTile * N)
                continue;
            const int col = e % kFftTile, p = e / kFftTile;
            const int m1 = m.nyquist ? m.outer0 + col : m.outer0;
            const int m2 = m.nyquist ? m.m2 : m.m2 + col;
            const int m0 = bitrev(p, LOG2N);
            const double h = hv[k];
            T2 v = buf[col][p];
            if (EV && h != 0.0) {
                const int h0 = (m0 <= a.K0 / 2) ? m0 : m0 - a.K0;
                const int h1 = (m1 <= a.K1 / 2) ? m1 : m1 - a.K1;
                const int h2 = (m2 <= a.K2 / 2) ? m2 : m2 - a.K2;
                const double mx = h0 * a.ic0, my = h1 * a.ic1, mz = h2 * a.ic2;
                const double msq = mx * mx + my * my + mz * mz;
                const double re = static_cast<double>(v.x), im = static_cast<double>(v.y);
                const double q2 = re * re + im * im;
                const double vt = -2.0 * (1.0 / msq + a.pi2_a2);
                const double ug = h * q2;
                const bool pair = m2 != 0 && 2 * m2 != a.K2;
                const double w = pair ? 2.0 : 1.0;
                const double sx = (2 * m0 == a.K0) ? 1.0 : -1.0;
                const double sy = (2 * m1 == a.K1) ? 1.0 : -1.0;
                const double oxy = pair ? 1.0 + sx * sy : 1.0;
                const double oxz = pair ? 1.0 - sx : 1.0;
                const double oyz = pair ? 1.0 - sy : 1.0;
                acc[0] += w * ug;
                acc[1] += w * ug * (1.0 + vt * mx * mx);
                acc[2] += w * ug * (1.0 + vt * my * my);
                acc[3] += w * ug * (1.0 + vt * mz * mz);
                acc[4] += oxy * ug * (vt * mx * my);
                acc[5] += oxz * ug * (vt * mx * mz);
                acc[6] += oyz * ug * (vt * my * mz);
            }
            v.x *= static_cast<T>(h);
            v.y *= static_cast<T>(h);
            buf[col][p] = v;
        }
        __syncthreads();
        if (col_w < kFftTile)
            dit<T2, LOG2N, true>(buf[col_w], tw, 1, t, NT);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = threadIdx.x + k * kFftTileThreads;
            if (e < kFftTile * N)
                c[m.base + (e % kFftTile) * m.cs + (e / kFftTile) * ls] = buf[e % kFftTile][e / kFftTile];
        }
    }
    if (EV)
        block_partials<kFftTileThreads>(acc, a.partials, blockIdx.x);
}

} // namespace pmefft
} // namespace atvh
