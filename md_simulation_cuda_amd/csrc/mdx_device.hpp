// Device-side building blocks of the gfx950 LJ step: pair parameters, the
// canonical cell coordinate, compensated sums and deterministic grid
// reductions.  Included by mdx_kernels.hip only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdx {

constexpr int kBlock = 256;          // threads per workgroup for every kernel
constexpr int kWaves = kBlock / 64;  // wave64
constexpr int kScanChunk = 1024;     // cells per workgroup in the two-pass scan

// Type class of a particle: 0 (type 0), 1 (type 1), 2 (any other type).  Pair
// parameters follow src/md_cuda_common.cu:220-229: (0,0)->AA, (1,1)->BB,
// everything else -> AB, i.e. index = (ci == cj) ? ci : 2 with AB stored at 2.
__host__ __device__ inline int pair_index(int ci, int cj) { return ci == cj ? ci : 2; }

// Per type-pair constants, index 0 = AA, 1 = BB, 2 = AB.
struct PairTable {
    float rc2f[3], sig2f[3], eps24f[3], eps4f[3];
    double rc2[3], sig2[3], eps24[3], eps4[3];
    double vc[3];          // eps * 4 (c^-12 - c^-6): energy shift
    double inv_m[3];       // 1 / mass by class (class 0 -> mass_a, 1,2 -> mass_b)
    double m[3];
    // FP32 neighbour-list pair in powers of 1/r^2: F/r = r^-2 r^-6 (fa r^-6 -
    // fb) with fa = 48 eps sigma^12, fb = 24 eps sigma^6; u = r^-6 (ua r^-6 -
    // ub) with ua = 4 eps sigma^12, ub = 4 eps sigma^6 (rounded from FP64)
    float fa[3], fb[3], ua[3], ub[3];
    float r2_guard;        // FP32 pairs closer than this are evaluated in FP64
    float pad;
    // compensated FP32 form: pairs with r^2 below (kCoreRatio sigma_ij)^2,
    // deep in the repulsive core, are evaluated in FP64 from their exact
    // fixed-point displacement
    float rcore2f[3];
    float pad2;
};

// Radius, in units of sigma_ij, inside which the compensated FP32 form
// evaluates a pair in FP64.  The displacement is exact (fixed point) up to
// its rounding to FP32 (half an ulp, ~3e-8 sigma at r ~ sigma); inside r ~
// sigma the force grows as r^-13 and that rounding moves it by up to ~1e-4
// epsilon/sigma.  At N = 1M (liquid, density 0.75, T 0.6) the largest
// per-particle force error max_i |dF_i| / rms |F| is 8.4e-6 with no FP64
// pairs, 6.4e-6 with those inside 0.95 sigma (about 0.002 of the ~15
// in-cutoff partners of a particle, so most waves have none), 4.2e-6 inside
// 1.0 sigma (0.09 per particle, 6% slower list pass).
constexpr double kCoreRatio = 0.95;

// Periodic wrap of include/md_cuda_common.hpp:20-29: fmod(x, L), plus L when
// negative.  The common cases return exactly what fmod returns (x in [0, L):
// x; x in [L, 2L): x - L, exact by Sterbenz; x in (-L, 0): x, then x + L)
// without calling the slow general fmod.
__host__ __device__ inline double pbc_wrap(double x, double L) {
    if (L <= 0.0) return x;
    if (x >= 0.0) {
        if (x < L) return x;
        if (x < 2.0 * L) return x - L;
    } else if (x > -L) {
        return x + L;
    }
    x = fmod(x, L);
    if (x < 0.0) x += L;
    return x;
}

// Workgroup remap for the 8 XCDs of MI355X: hardware places workgroup i on
// XCD i % 8, so consecutive ids land on different L2s.  The returned logical
// id gives each XCD one contiguous range of logical workgroups (spatially
// adjacent cells), keeping neighbour-cell reads in the same L2.  A bijection
// on [0, nb).
__device__ inline int xcd_block(int i, int nb) {
    constexpr int kXcd = 8;
    const int q = nb / kXcd, rem = nb % kXcd;
    const int x = i % kXcd, k = i / kXcd;
    return x * q + (x < rem ? x : rem) + k;
}

// 1/x in FP32: hardware reciprocal refined by one Newton step (error about
// 1 ulp, against ~10 instructions for the IEEE division).
__device__ inline float fast_rcp(float x) {
    float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_fmaf(r, e, r);
}

// Canonical cell coordinate: floor(x * n / L) from the FP64 position,
// clamped to [0, n-1].  The oracle evaluates the identical expression.
__host__ __device__ inline int cell_coord(double x, double L, int n) {
    const double q = (x * static_cast<double>(n)) / L;
    if (!(q >= 0.0)) return 0;
    if (q >= static_cast<double>(n)) return n - 1;
    return static_cast<int>(q);
}

// Geometry of one cell grid.
//
// The FP32 forms keep each particle's position relative to its cell's centre
// as fixed-point int32, coordinate * 2^qk.  A pair's displacement is then the
// difference of two such images less the neighbour cell's offset (wq, hq),
// formed exactly in integer arithmetic; only its conversion to FP32 rounds.
// qk is the largest exponent for which 5 max(w, h) * 2^qk < 2^31: images lie
// within w / 2 (plus the drift since the last build) of the centre, and the
// widest displacement formed, towards a cell two columns away, stays below
// 5 max(w, h), so no difference wraps.
struct Grid {
    double Lx, Ly;
    double w, h;       // cell sides
    float Lxf, Lyf, wf, hf;
    int ncx, ncy;
    int periodic_y;
    int small_x;       // cell side < rc_max in x: per-pair minimum image
    int small_y;
    int qk;            // fixed-point exponent of the FP32 forms' images
    int wq, hq;        // cell sides in fixed point (rounded)
    int Lxq, Lyq;      // box sides in fixed point, for the minimum image on small grids
                       // (there Lx = 3 w fits; elsewhere clamped and unused)
    float q2f;         // 2^(2 qk): r^2 from fixed-point units to FP32 units of sigma^2
    float qinvf;       // 2^-qk
    double qs, qinv;   // 2^qk, 2^-qk
};

// Fixed-point image coordinate: d * 2^qk rounded to the nearest integer,
// clamped to the int32 range (a particle far outside a non-periodic box).
__host__ __device__ inline int to_fixed(double d, double qs) {
    double v = rint(d * qs);
    v = v > 2147483520.0 ? 2147483520.0 : (v < -2147483520.0 ? -2147483520.0 : v);
    return static_cast<int>(v);
}

// Error-free transformation a + b = s + e (Knuth TwoSum), used to carry the
// rounding error of the FP32 force sums.
struct Comp32 {
    float s, c;
    __device__ inline void add(float t) {
        const float snew = s + t;
        const float bp = snew - s;
        const float err = (s - (snew - bp)) + (t - bp);
        s = snew;
        c += err;
    }
    __device__ inline double value() const { return static_cast<double>(s) + static_cast<double>(c); }
};

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Workgroup part of a deterministic two-kernel reduction: the workgroup
// reduces NV doubles in a fixed tree and thread 0 stores one partial per
// component at partials[k * stride + gblk].  Component k is a max when bit k of
// maxmask is set, a sum otherwise.  A later kernel (slot_reduce) combines a
// slot's partials in index order; the kernel boundary orders the stores, so no
// device-scope fence or atomic is needed.  Requires blockDim.x == 64 * NW.
template <int NV, int NW = kWaves>
__device__ void block_partials(double (&v)[NV], unsigned maxmask, double *partials, int stride, int gblk) {
    __shared__ double sh[NV][NW];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const bool mx = (maxmask >> k) & 1u;
        const double r = mx ? wave_max(v[k]) : wave_sum(v[k]);
        if (lane == 0) sh[k][wid] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const bool mx = (maxmask >> k) & 1u;
            double b = sh[k][0];
            for (int w = 1; w < NW; ++w) b = mx ? fmax(b, sh[k][w]) : b + sh[k][w];
            partials[k * stride + gblk] = b;
        }
    }
}

// Second part: one workgroup combines nb partials (partials[k * stride + base
// + i], i < nb) in a fixed order.  Every thread returns the totals.  The first
// kBlock threads combine, so a larger workgroup gives the same order (and the
// same totals).
template <int NV>
__device__ void slot_reduce(unsigned maxmask, const double *partials, int stride, int base, int nb,
                            double (&out)[NV]) {
    __shared__ double sh[NV][kWaves];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    // thread t combines partials t, t + kBlock, t + 2 kBlock, ... in that
    // order; eight loads per value are in flight before the first add
    constexpr int kU = 8;
    double a[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) a[k] = ((maxmask >> k) & 1u) ? -1.0e308 : 0.0;
    int i = threadIdx.x < kBlock ? static_cast<int>(threadIdx.x) : nb;
    for (; i + (kU - 1) * kBlock < nb; i += kU * kBlock) {
        double q[NV][kU];
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int u = 0; u < kU; ++u) q[k][u] = partials[k * stride + base + i + u * kBlock];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const bool mx = (maxmask >> k) & 1u;
#pragma unroll
            for (int u = 0; u < kU; ++u) a[k] = mx ? fmax(a[k], q[k][u]) : a[k] + q[k][u];
        }
    }
    for (; i < nb; i += kBlock) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const bool mx = (maxmask >> k) & 1u;
            const double q = partials[k * stride + base + i];
            a[k] = mx ? fmax(a[k], q) : a[k] + q;
        }
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const bool mx = (maxmask >> k) & 1u;
        const double r = mx ? wave_max(a[k]) : wave_sum(a[k]);
        if (lane == 0 && wid < kWaves) sh[k][wid] = r;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const bool mx = (maxmask >> k) & 1u;
        double b = sh[k][0];
        for (int w = 1; w < kWaves; ++w) b = mx ? fmax(b, sh[k][w]) : b + sh[k][w];
        out[k] = b;
    }
}

}  // namespace mdx
