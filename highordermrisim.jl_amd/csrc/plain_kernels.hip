// plain_kernels.hip -- per-block drop-in kernels in the plan's precision
// (f32 or f64): the reference GPU algorithm restated for CDNA4.
//
//   k_plain_precession   src/simulation/kernels/HighOrderGPU.jl:63-122
//   k_plain_excitation   src/simulation/kernels/HighOrderGPU.jl:181-221
//   k_coil_partials/     src/simulation/kernels/CoilReductionGPU.jl:16-100
//   k_coil_finalize      (deterministic two-stage coil reduction, no atomics)
//
// One lane per spin, node loop in registers; node data is wave-uniform and is
// read through the scalar cache.  The field-order mask is a template
// parameter so each mask compiles to straight-line code without the disabled
// terms, as the reference does with Val(ORDERS).
#include "hom_internal.h"

#include <algorithm>
#include <mutex>

namespace hom {
namespace {

template <typename T>
struct alignas(2 * sizeof(T)) Cx {
    T re, im;
};

template <typename T>
__device__ inline Cx<T> cmul(Cx<T> a, Cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ inline Cx<T> cscale(Cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T>
__device__ inline Cx<T> cconj(Cx<T> a) { return {a.re, -a.im}; }

__device__ inline void dsincos(double a, double *s, double *c) { sincos(a, s, c); }
__device__ inline void dsincos(float a, float *s, float *c) { sincosf(a, s, c); }
__device__ inline double dexp(double a) { return exp(a); }
__device__ inline float dexp(float a) { return expf(a); }
__device__ inline double dsqrt(double a) { return sqrt(a); }
__device__ inline float dsqrt(float a) { return sqrtf(a); }
template <typename T> __device__ inline T eps_of();
template <> __device__ inline double eps_of<double>() { return 2.220446049250313e-16; }
template <> __device__ inline float eps_of<float>() { return 1.1920929e-7f; }

template <typename T>
__device__ inline Cx<T> cis(T a) {
    T s, c;
    dsincos(a, &s, &c);
    return {c, s};
}

// Masked field of one spin at one node (HighOrderGPU.jl:73-93).
//   MASK bit 0..3: ho0..ho3, bit 5: dw_precession.
template <typename T, int MASK>
__device__ inline T field_at(const T (&b)[16], const T *__restrict__ coef, int64_t nn,
                             const T *__restrict__ gx, const T *__restrict__ gy,
                             const T *__restrict__ gz, int64_t n, T dw_term) {
    const T bz0 = (MASK & HOM_HO0) ? coef[n] : T(0);
    T bz1;
    if (MASK & HOM_HO1)
        bz1 = b[1] * coef[nn + n] + b[2] * coef[2 * nn + n] + b[3] * coef[3 * nn + n];
    else
        bz1 = b[1] * gx[n] + b[2] * gy[n] + b[3] * gz[n];
    T bz2 = T(0), bz3 = T(0);
    if (MASK & HOM_HO2)
        bz2 = b[4] * coef[4 * nn + n] + b[5] * coef[5 * nn + n] + b[6] * coef[6 * nn + n] +
              b[7] * coef[7 * nn + n] + b[8] * coef[8 * nn + n];
    if (MASK & HOM_HO3)
        bz3 = b[9] * coef[9 * nn + n] + b[10] * coef[10 * nn + n] + b[11] * coef[11 * nn + n] +
              b[12] * coef[12 * nn + n] + b[13] * coef[13 * nn + n] + b[14] * coef[14 * nn + n] +
              b[15] * coef[15 * nn + n];
    const T bz = bz0 + bz1 + bz2 + bz3;
    return (MASK & HOM_DW_PRECESSION) ? bz + dw_term : bz;
}

// Unnormalised real solid harmonics through third order (HighOrderGPU.jl:63-71).
template <typename T>
__device__ inline void basis16(T x, T y, T z, T (&b)[16]) {
    const T x2 = x * x, y2 = y * y, z2 = z * z;
    const T r2 = x2 + y2 + z2;
    b[0] = T(1); b[1] = x; b[2] = y; b[3] = z;
    b[4] = x * y; b[5] = z * y; b[6] = T(3) * z2 - r2; b[7] = x * z; b[8] = x2 - y2;
    b[9] = T(3) * y * x2 - y * y2; b[10] = x * z * y; b[11] = (T(5) * z2 - r2) * y;
    b[12] = T(5) * z * z2 - T(3) * z * r2; b[13] = (T(5) * z2 - r2) * x;
    b[14] = x2 * z - y2 * z; b[15] = x * x2 - T(3) * x * y2;
}

template <typename T, int MASK>
__global__ __launch_bounds__(256) void k_plain_precession(
    int64_t ns, const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
    const T *__restrict__ rho, const T *__restrict__ T1, const T *__restrict__ T2,
    const T *__restrict__ dw, const T *__restrict__ coef, int64_t nn, const T *__restrict__ gx,
    const T *__restrict__ gy, const T *__restrict__ gz, const T *__restrict__ dt,
    const uint8_t *__restrict__ adc, int64_t first, int64_t last, T kphase, T inv2pig,
    Cx<T> *__restrict__ mxy, T *__restrict__ mz, Cx<T> *__restrict__ states, int64_t ld) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    T b[16];
    basis16(x[s], y[s], z[s], b);
    const T dw_term = dw[s] * inv2pig;
    const T t2 = T2[s];
    T bprev = field_at<T, MASK>(b, coef, nn, gx, gy, gz, first, dw_term);
    const Cx<T> initial = mxy[s];
    T phi = T(0), t = T(0);
    int64_t sample = 0;
    for (int64_t n = first + 1; n <= last; ++n) {
        const T d = dt[n - 1];
        t += d;
        const T bnext = field_at<T, MASK>(b, coef, nn, gx, gy, gz, n, dw_term);
        phi += kphase * (bprev + bnext) * d;
        if (adc[n]) {
            states[s + sample * ld] = cmul(cscale(initial, dexp(-t / t2)), cis(phi));
            ++sample;
        }
        bprev = bnext;
    }
    mxy[s] = cmul(cscale(initial, dexp(-t / t2)), cis(phi));
    const T e1 = dexp(-t / T1[s]);
    mz[s] = mz[s] * e1 + rho[s] * (T(1) - e1);
}

// Hard-pulse spinor update (HighOrderGPU.jl:181-192).
template <typename T>
__device__ inline void rotate_spin(Cx<T> &m, T &mzv, Cx<T> b1, T bz, T dt, T k2pig) {
    T B = dsqrt(b1.re * b1.re + b1.im * b1.im + bz * bz);
    B = (B == T(0)) ? eps_of<T>() : B;
    const T phi = k2pig * (B * dt);
    T s, c;
    dsincos(phi / T(2), &s, &c);
    const Cx<T> a = {c, -(bz / B) * s};
    const Cx<T> nb = {b1.re / B * s, b1.im / B * s};
    const Cx<T> beta = {nb.im, -nb.re};  // -i * (B1/B) * s
    const Cx<T> ca = cconj(a);
    const Cx<T> t1 = cscale(cmul(ca, beta), T(2) * mzv);
    const Cx<T> t2 = cmul(cmul(ca, ca), m);
    const Cx<T> t3 = cmul(cmul(beta, beta), cconj(m));
    const Cx<T> ab = cmul(cmul(a, beta), cconj(m));
    const T na = a.re * a.re + a.im * a.im, nbeta = beta.re * beta.re + beta.im * beta.im;
    m = {t1.re + t2.re - t3.re, t1.im + t2.im - t3.im};
    mzv = (na - nbeta) * mzv - T(2) * ab.re;
}

template <typename T, bool DWX>
__global__ __launch_bounds__(256) void k_plain_excitation(
    int64_t ns, const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
    const T *__restrict__ rho, const T *__restrict__ T1, const T *__restrict__ T2,
    const T *__restrict__ dw, const T *__restrict__ gx, const T *__restrict__ gy,
    const T *__restrict__ gz, const Cx<T> *__restrict__ b1, const T *__restrict__ df,
    const T *__restrict__ psi, const T *__restrict__ dt, const uint8_t *__restrict__ adc,
    int64_t first, int64_t last, T k2pig, T inv2pig, T invg, Cx<T> *__restrict__ mxy,
    T *__restrict__ mz, Cx<T> *__restrict__ states, int64_t ld) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    Cx<T> m = cmul(mxy[s], cis(-psi[first]));
    T mzv = mz[s];
    const T off = DWX ? dw[s] * inv2pig : T(0);
    const T xs = x[s], ys = y[s], zs = z[s], t1s = T1[s], t2s = T2[s], rs = rho[s];
    int64_t sample = 0;
    for (int64_t n = first; n < last; ++n) {
        const T d = dt[n];
        const T bz = gx[n] * xs + gy[n] * ys + gz[n] * zs + off - df[n] * invg;
        rotate_spin(m, mzv, b1[n], bz, d, k2pig);
        m = cscale(m, dexp(-d / t2s));
        const T e1 = dexp(-d / t1s);
        mzv = mzv * e1 + rs * (T(1) - e1);
        if (adc[n + 1]) {
            states[s + sample * ld] = cmul(m, cis(psi[n + 1]));
            ++sample;
        }
    }
    mxy[s] = cmul(m, cis(psi[last]));
    mz[s] = mzv;
}

// Deterministic coil reduction, stage 1: the complex GEMM statesᵀ·csm on
// the matrix cores as a real GEMM: K runs over (spin, re|im), the A operand
// is the state's real or imaginary part, the B operand (re, im; -im, re) of
// the sensitivity, so the 16 output columns of one MFMA tile are 8 coils x
// (re, im).  One workgroup per (spin share, 32-sample x 32-coil channel
// tile); the spins go in chunks of kCoilChunk<T> spins (256 bytes of a row),
// and workgroup g of G takes chunks g, g + G, g + 2G, ... so that the
// workgroups in flight read neighbouring runs of every state and sensitivity
// row (owning contiguous spin tiles instead ran 8-19 % slower).
// G = ceil(Ns / kSpinTile): 512 workgroups for 1M spins, the fastest measured
// (of 256-2048 for FP64, 512-4096 for FP32).  Wave w owns sample tile w & 1
// and coils 16 (w >> 1) .. +15 (two 16x16 accumulators); chunks pass through
// LDS spin-major, rows padded by one
// element (conflict-free transposed stores and fragment reads),
// double-buffered, with the global loads two chunks ahead of the compute.
// The partial of workgroup g sums its chunks in a fixed order: the result is
// deterministic.
constexpr int64_t kSpinTile = 2048;
constexpr int64_t kSpinTileMin = kSpinTile;
// Spins per chunk: 256 bytes of one state or sensitivity row (16 FP64 or 32
// FP32 complex values).  Measured on a 32-sample block of 1M spins and 32
// coils: FP64 259 us with 16-spin and 278 us with 32-spin chunks, FP32 157 us
// with 16 and 140 us with 32.
template <typename T>
constexpr int kCoilChunk = 256 / (int)sizeof(Cx<T>);

template <typename T>
struct CoilMfma;
template <>
struct CoilMfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    __device__ static acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // output row of accumulator element v in lane l (column l % 16)
    __device__ static int row(int l, int v) { return (l >> 4) + 4 * v; }
};
template <>
struct CoilMfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    __device__ static acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int l, int v) { return 4 * (l >> 4) + v; }
};

template <typename T>
__global__ __launch_bounds__(256) void k_coil_partials(
    int64_t ns, int64_t nadc, int64_t nc, const Cx<T> *__restrict__ states, int64_t lds,
    const Cx<T> *__restrict__ csm, int64_t ldc, Cx<T> *__restrict__ partials, int64_t nch) {
    using M = CoilMfma<T>;
    constexpr int kC = kCoilChunk<T>;
    __shared__ Cx<T> st[2][kC][33];  // [buffer][spin][sample]
    __shared__ Cx<T> cs[2][kC][33];  // [buffer][spin][coil]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t tile = blockIdx.x;
    const int64_t ktiles = (nadc + 31) / 32;
    const int64_t k0 = (blockIdx.y % ktiles) * 32, c0 = (blockIdx.y / ktiles) * 32;
    // chunk i of this workgroup: global chunk tile + i * gridDim.x
    const int64_t nchunks = (ns + kC - 1) / kC, stride = gridDim.x;
    const int64_t nck = (nchunks - 1 - tile) / stride + 1;  // tile < nchunks
    auto chunk0 = [&](int64_t i) { return (tile + i * stride) * kC; };
    const int64_t s_hi = ns;

    // staging: element e = tid + 256 j of a chunk is spin e % kC, sample/coil e / kC;
    // two register sets (chunks i + 1 and i + 2 in flight while chunk i is in use)
    constexpr int kJ = kC * 32 / 256;  // staged elements per thread and array
    struct Stage {
        Cx<T> s[kJ], c[kJ];
    };
    // Loads are unconditional, from clamped in-bounds addresses (ns, nadc and
    // nc are at least 1 here), and the bounds apply as selects at the store:
    // with no branch around them the compiler waits only for the older
    // register set before a store, not for every load in flight.
    auto load = [&](Stage &r, int64_t i) {
        const int64_t sc = chunk0(i);
#pragma unroll
        for (int j = 0; j < kJ; ++j) {
            const int e = tid + 256 * j, sl = e % kC, q = e / kC;
            const int64_t sp = std::min<int64_t>(sc + sl, s_hi - 1);
            r.s[j] = states[sp + std::min<int64_t>(k0 + q, nadc - 1) * lds];
            r.c[j] = csm[sp + std::min<int64_t>(c0 + q, nc - 1) * ldc];
        }
    };
    auto store = [&](const Stage &r, int b, int64_t i) {
        const int64_t sc = chunk0(i);
#pragma unroll
        for (int j = 0; j < kJ; ++j) {
            const int e = tid + 256 * j, sl = e % kC, q = e / kC;
            const bool in = i < nck && sc + sl < s_hi;
            st[b][sl][q] = (in && k0 + q < nadc) ? r.s[j] : Cx<T>{};
            cs[b][sl][q] = (in && c0 + q < nc) ? r.c[j] : Cx<T>{};
        }
    };

    // fragment roles of this lane: K index kk = lane / 16 -> spin kk / 2 of the
    // step, part kk % 2; A row (sample) and B column (coil, re|im) lane % 16
    const int mt = w & 1, nh = w >> 1;
    const int kk = lane >> 4, n = lane & 15;
    const int part = kk & 1, q = n & 1;
    const bool b_re = part == q;                       // (re, im; -im, re)
    const T b_sign = (part == 1 && q == 0) ? T(-1) : T(1);
    const int a_col = 16 * mt + n;
    const int b_col0 = 16 * nh + (n >> 1);
    typename M::acc_t acc[2] = {};
    auto compute = [&](int b) {
#pragma unroll
        for (int ks = 0; ks < kC / 2; ++ks) {
            const int sl = 2 * ks + (kk >> 1);
            const Cx<T> av = st[b][sl][a_col];
            const T a = part ? av.im : av.re;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const Cx<T> bv = cs[b][sl][b_col0 + 8 * j];
                acc[j] = M::mma(a, (b_re ? bv.re : bv.im) * b_sign, acc[j]);
            }
        }
    };

    // chunks in pairs: even chunks in LDS buffer 0 via register set ra, odd
    // chunks in buffer 1 via rb.  Stores are unconditional as well: one past
    // the last chunk writes zeros into a buffer that is not read again.
    Stage ra, rb;
    load(ra, 0);
    load(rb, 1);
    store(ra, 0, 0);
    __syncthreads();
    // sched_barrier: keep each chunk's loads ahead of the compute (the
    // scheduler otherwise sinks them to the end of it)
    for (int64_t i = 0; i < nck; i += 2) {
        load(ra, i + 2);
        __builtin_amdgcn_sched_barrier(0);
        compute(0);
        store(rb, 1, i + 1);
        __syncthreads();
        if (i + 1 >= nck) break;
        load(rb, i + 3);
        __builtin_amdgcn_sched_barrier(0);
        compute(1);
        store(ra, 0, i + 2);
        __syncthreads();
    }
    // channel-major partials: the tiles of one (sample, coil) channel are contiguous
    const int64_t ntiles = gridDim.x;
    T *out = reinterpret_cast<T *>(partials);
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t k = k0 + 16 * mt + M::row(lane, v);
            const int64_t c = c0 + 16 * nh + 8 * j + (n >> 1);
            if (k < nadc && c < nc) out[((k + c * nadc) * ntiles + tile) * 2 + q] = acc[j][v];
        }
}

// Stage 2, as the reference's finalize (CoilReductionGPU.jl:44-69): one
// workgroup per (sample, coil) channel; lane l sums tiles l, l + 256, ... and
// a fixed tree in LDS combines the lanes, then the channel is written to the
// (possibly strided) output.
template <typename T>
__global__ __launch_bounds__(256) void k_coil_finalize(int64_t ntiles, int64_t nadc,
                                                      const Cx<T> *__restrict__ partials,
                                                      Cx<T> *__restrict__ out, int64_t ldo) {
    __shared__ Cx<T> red[256];
    const int64_t ch = blockIdx.x;
    const int lane = threadIdx.x;
    const Cx<T> *row = partials + ch * ntiles;
    Cx<T> v = {};
    for (int64_t t = lane; t < ntiles; t += 256) {
        const Cx<T> q = row[t];
        v.re += q.re;
        v.im += q.im;
    }
    red[lane] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (lane < off) {
            red[lane].re += red[lane + off].re;
            red[lane].im += red[lane + off].im;
        }
        __syncthreads();
    }
    if (lane == 0) out[ch % nadc + (ch / nadc) * ldo] = red[0];
}

// ---- fused precession + coil reception (coil modes "fused_warp" and
// "fused_tiled"; FusedCoilGPU.jl:70-133) ----
// One thread per spin and 256 spins per workgroup, the node loop of
// k_plain_precession; the ADC states of a workgroup go to LDS instead of HBM,
// kFusedTile samples at a time, and each full (or final) tile is multiplied
// by the sensitivities on the matrix cores right there (same real-GEMM form
// as k_coil_partials: K = (spin, re|im), 16 columns = 8 coils x (re, im)).
// The units of a tile are (8-coil block, spin part): with fewer than 17
// coils the 256 spins split into 4 / ceil(nc / 8) parts so that every wave has
// a unit, and each part writes its own partial row.  Every partial (row,
// sample, coil) is written exactly once per block and k_coil_finalize sums
// the rows in a fixed order, so the result is deterministic.  The state
// buffer is not used: a fused plan only sizes it for ADC during RF.
constexpr int kFusedSpins = 256;
constexpr int kFusedTile = 16;               // samples per tile: the MFMA's M
constexpr int kFusedRow = kFusedTile + 1;    // LDS row stride (complex), padded

inline int fused_spin_parts(int64_t nc) {
    const int64_t ncb = (nc + 7) / 8;
    return ncb >= 3 ? 1 : (int)(4 / ncb);
}

template <typename T>
__device__ inline void fused_tile_gemm(const Cx<T> *st, int64_t g, int64_t ns,
                                       const Cx<T> *__restrict__ csm, int64_t ldc, int64_t nc,
                                       int parts, int64_t k0, int cnt, int64_t nadc,
                                       int64_t nrows, T *__restrict__ out) {
    using M = CoilMfma<T>;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int kk = lane >> 4, n = lane & 15;
    const int part = kk & 1, q = n & 1;
    const bool b_re = part == q;  // B = (re, im; -im, re) of the sensitivity
    const T b_sign = (part == 1 && q == 0) ? T(-1) : T(1);
    const int ncb = (int)((nc + 7) / 8);
    const int span = kFusedSpins / parts;
    const int nks = span / 2;  // K steps of 2 spins: 128, 64 or 32
    for (int u = w; u < ncb * parts; u += 4) {
        const int cb = u % ncb, sp = u / ncb;
        const int64_t c = (int64_t)cb * 8 + (n >> 1);
        const bool c_ok = c < nc;
        // the one component of the sensitivity this lane's B entry needs
        const T *colv = reinterpret_cast<const T *>(csm + (c_ok ? c : nc - 1) * ldc) + (b_re ? 0 : 1);
        const T bs = c_ok ? b_sign : T(0);
        const int sl0 = sp * span + (kk >> 1);
        const int64_t sg0 = g * kFusedSpins + sl0;
        auto bval = [&](int ks) {
            const int64_t sg = sg0 + 2 * ks;
            return sg < ns ? colv[2 * sg] * bs : T(0);
        };
        typename M::acc_t acc = {};
        auto step = [&](int ks, T b) {
            const Cx<T> av = st[(sl0 + 2 * ks) * kFusedRow + n];
            acc = M::mma(part ? av.im : av.re, b, acc);
        };
        if constexpr (sizeof(T) == 8) {
            // FP64: sensitivities in batches of kB K steps, the next batch
            // loading while the current one feeds the MFMAs (measured on 1M
            // spins x 32 coils: 3.55e10 against 2.87e10 spin-samples/s with
            // an unrolled loop; in FP32 the unrolled loop is faster, 6.2e10
            // against 5.4e10)
            constexpr int kB = 16;
            T bq[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j) bq[j] = bval(j);
            for (int ks0 = 0; ks0 < nks; ks0 += kB) {
                T bn[kB];
#pragma unroll
                for (int j = 0; j < kB; ++j) bn[j] = ks0 + kB + j < nks ? bval(ks0 + kB + j) : T(0);
#pragma unroll
                for (int j = 0; j < kB; ++j) step(ks0 + j, bq[j]);
#pragma unroll
                for (int j = 0; j < kB; ++j) bq[j] = bn[j];
            }
        } else {
            // whole complex loads (8 bytes), the component selected after
            // the load: 6.2e10 against 5.5e10 with 4-byte component loads
            const Cx<T> *col = csm + (c_ok ? c : nc - 1) * ldc;
#pragma unroll 8
            for (int ks = 0; ks < nks; ++ks) {
                const int64_t sg = sg0 + 2 * ks;
                const bool s_ok = sg < ns;
                const Cx<T> bv = col[s_ok ? sg : ns - 1];
                step(ks, s_ok ? (b_re ? bv.re : bv.im) * bs : T(0));
            }
        }
        if (c_ok) {
            const int64_t row = g * parts + sp;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int k = M::row(lane, v);
                if (k < cnt) out[((k0 + k + c * nadc) * nrows + row) * 2 + q] = acc[v];
            }
        }
    }
}

template <typename T, int MASK>
__global__ __launch_bounds__(256) void k_fused_precession(
    int64_t ns, const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
    const T *__restrict__ rho, const T *__restrict__ T1, const T *__restrict__ T2,
    const T *__restrict__ dw, const T *__restrict__ coef, int64_t nn, const T *__restrict__ gx,
    const T *__restrict__ gy, const T *__restrict__ gz, const T *__restrict__ dt,
    const uint8_t *__restrict__ adc, int64_t first, int64_t last, T kphase, T inv2pig,
    Cx<T> *__restrict__ mxy, T *__restrict__ mz, const Cx<T> *__restrict__ csm, int64_t ldc,
    int64_t nc, int64_t nadc, int parts, int64_t nrows, T *__restrict__ partials) {
    __shared__ Cx<T> st[kFusedSpins * kFusedRow];  // [spin][sample]
    const int64_t g = blockIdx.x;
    const int64_t s = g * kFusedSpins + threadIdx.x;
    // Every thread runs the node loop (the tile barriers are workgroup-wide);
    // spins past ns compute on spin ns - 1 (ns >= 1) and store zero states.
    const bool active = s < ns;
    const int64_t si = active ? s : ns - 1;
    T b[16];
    basis16(x[si], y[si], z[si], b);
    const T dw_term = dw[si] * inv2pig;
    const T t2 = T2[si];
    T bprev = field_at<T, MASK>(b, coef, nn, gx, gy, gz, first, dw_term);
    const Cx<T> initial = mxy[si];
    Cx<T> *mine = st + threadIdx.x * kFusedRow;
    T phi = T(0), t = T(0);
    int cnt = 0;
    int64_t k0 = 0;
    for (int64_t n = first + 1; n <= last; ++n) {
        const T d = dt[n - 1];
        t += d;
        const T bnext = field_at<T, MASK>(b, coef, nn, gx, gy, gz, n, dw_term);
        phi += kphase * (bprev + bnext) * d;
        if (adc[n]) {  // the same for every thread: the barriers are uniform
            mine[cnt] = active ? cmul(cscale(initial, dexp(-t / t2)), cis(phi)) : Cx<T>{};
            if (++cnt == kFusedTile) {
                __syncthreads();
                fused_tile_gemm<T>(st, g, ns, csm, ldc, nc, parts, k0, cnt, nadc, nrows, partials);
                __syncthreads();  // the tile is read before the next states overwrite it
                k0 += kFusedTile;
                cnt = 0;
            }
        }
        bprev = bnext;
    }
    if (cnt > 0) {
        __syncthreads();
        fused_tile_gemm<T>(st, g, ns, csm, ldc, nc, parts, k0, cnt, nadc, nrows, partials);
    }
    if (active) {
        mxy[s] = cmul(cscale(initial, dexp(-t / t2)), cis(phi));
        const T e1 = dexp(-t / T1[s]);
        mz[s] = mz[s] * e1 + rho[s] * (T(1) - e1);
    }
}

template <typename T>
__global__ void k_fill_zero(Cx<T> *out, int64_t rows, int64_t cols, int64_t ld) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    out[(i % rows) + (i / rows) * ld] = Cx<T>{};
}

template <typename T>
__global__ void k_reset_state(int64_t ns, const T *rho, Cx<T> *mxy, T *mz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    mxy[i] = Cx<T>{};
    mz[i] = rho[i];
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

template <typename T>
hipError_t precession_t(hom_plan *p, int64_t first, int64_t last, int flags) {
    const int mask = flags & (HOM_HO0 | HOM_HO1 | HOM_HO2 | HOM_HO3 | HOM_DW_PRECESSION);
    const auto &d = p->d;
    const T kphase = T(-M_PI * p->gamma);
    const T inv2pig = T(1) / T(kTwoPi * p->gamma);
    if (p->ns == 0) return hipSuccess;
    dim3 grid(grid_for(p->ns, 256)), block(256);
#define HOM_CASE(M)                                                                           \
    case M:                                                                                   \
        hipLaunchKernelGGL((k_plain_precession<T, M>), grid, block, 0, p->stream, p->ns,      \
                           (const T *)d.x, (const T *)d.y, (const T *)d.z, (const T *)d.rho,  \
                           (const T *)d.T1, (const T *)d.T2, (const T *)d.dw,                 \
                           (const T *)d.coef, p->nn, (const T *)d.gx, (const T *)d.gy,        \
                           (const T *)d.gz, (const T *)d.dt, d.adc, first, last, kphase,      \
                           inv2pig, (Cx<T> *)d.mxy, (T *)d.mz, (Cx<T> *)d.states, p->ns);     \
        break;
    ScopedTimer timer(p, HOM_K_PLAIN_PRECESSION);
    switch (mask) {
        HOM_CASE(0) HOM_CASE(1) HOM_CASE(2) HOM_CASE(3) HOM_CASE(4) HOM_CASE(5) HOM_CASE(6)
        HOM_CASE(7) HOM_CASE(8) HOM_CASE(9) HOM_CASE(10) HOM_CASE(11) HOM_CASE(12)
        HOM_CASE(13) HOM_CASE(14) HOM_CASE(15) HOM_CASE(32) HOM_CASE(33) HOM_CASE(34)
        HOM_CASE(35) HOM_CASE(36) HOM_CASE(37) HOM_CASE(38) HOM_CASE(39) HOM_CASE(40)
        HOM_CASE(41) HOM_CASE(42) HOM_CASE(43) HOM_CASE(44) HOM_CASE(45) HOM_CASE(46)
        HOM_CASE(47)
        default: return hipErrorInvalidValue;
    }
#undef HOM_CASE
    return hipGetLastError();
}

template <typename T>
hipError_t fused_t(hom_plan *p, int64_t first, int64_t last, int flags, int64_t nadc, void *out,
                   int64_t ldo) {
    const int mask = flags & (HOM_HO0 | HOM_HO1 | HOM_HO2 | HOM_HO3 | HOM_DW_PRECESSION);
    const auto &d = p->d;
    if (p->ns == 0) {
        if (nadc > 0)
            hipLaunchKernelGGL((k_fill_zero<T>), dim3(grid_for(nadc * p->nc, 256)), dim3(256), 0,
                               p->stream, (Cx<T> *)out, nadc, p->nc, ldo);
        return hipGetLastError();
    }
    const T kphase = T(-M_PI * p->gamma);
    const T inv2pig = T(1) / T(kTwoPi * p->gamma);
    const int parts = fused_spin_parts(p->nc);
    const int64_t groups = (p->ns + kFusedSpins - 1) / kFusedSpins;
    const int64_t nrows = groups * parts;
    const int64_t ldc = p->csm_ld ? p->csm_ld : p->ns;
    dim3 grid((unsigned)groups), block(kFusedSpins);
#define HOM_CASE(M)                                                                           \
    case M:                                                                                   \
        hipLaunchKernelGGL((k_fused_precession<T, M>), grid, block, 0, p->stream, p->ns,      \
                           (const T *)d.x, (const T *)d.y, (const T *)d.z, (const T *)d.rho,  \
                           (const T *)d.T1, (const T *)d.T2, (const T *)d.dw,                 \
                           (const T *)d.coef, p->nn, (const T *)d.gx, (const T *)d.gy,        \
                           (const T *)d.gz, (const T *)d.dt, d.adc, first, last, kphase,      \
                           inv2pig, (Cx<T> *)d.mxy, (T *)d.mz, (const Cx<T> *)d.csm, ldc,     \
                           p->nc, nadc, parts, nrows, (T *)d.partials);                       \
        break;
    {
        ScopedTimer timer(p, HOM_K_FUSED_PRECESSION);
        switch (mask) {
            HOM_CASE(0) HOM_CASE(1) HOM_CASE(2) HOM_CASE(3) HOM_CASE(4) HOM_CASE(5) HOM_CASE(6)
            HOM_CASE(7) HOM_CASE(8) HOM_CASE(9) HOM_CASE(10) HOM_CASE(11) HOM_CASE(12)
            HOM_CASE(13) HOM_CASE(14) HOM_CASE(15) HOM_CASE(32) HOM_CASE(33) HOM_CASE(34)
            HOM_CASE(35) HOM_CASE(36) HOM_CASE(37) HOM_CASE(38) HOM_CASE(39) HOM_CASE(40)
            HOM_CASE(41) HOM_CASE(42) HOM_CASE(43) HOM_CASE(44) HOM_CASE(45) HOM_CASE(46)
            HOM_CASE(47)
            default: return hipErrorInvalidValue;
        }
    }
#undef HOM_CASE
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nadc == 0) return e;
    ScopedTimer timer(p, HOM_K_COIL_FINALIZE);
    hipLaunchKernelGGL((k_coil_finalize<T>), dim3((unsigned)(nadc * p->nc)), dim3(256), 0,
                       p->stream, nrows, nadc, (const Cx<T> *)d.partials, (Cx<T> *)out, ldo);
    return hipGetLastError();
}

template <typename T>
hipError_t excitation_t(hom_plan *p, int64_t first, int64_t last, int flags) {
    const auto &d = p->d;
    if (p->ns == 0) return hipSuccess;
    const T k2pig = T(-kTwoPi * p->gamma);
    const T inv2pig = T(1) / T(kTwoPi * p->gamma);
    const T invg = T(1) / T(p->gamma);
    dim3 grid(grid_for(p->ns, 256)), block(256);
    ScopedTimer timer(p, HOM_K_PLAIN_EXCITATION);
    if (flags & HOM_DW_EXCITATION)
        hipLaunchKernelGGL((k_plain_excitation<T, true>), grid, block, 0, p->stream, p->ns,
                           (const T *)d.x, (const T *)d.y, (const T *)d.z, (const T *)d.rho,
                           (const T *)d.T1, (const T *)d.T2, (const T *)d.dw, (const T *)d.gx,
                           (const T *)d.gy, (const T *)d.gz, (const Cx<T> *)d.b1,
                           (const T *)d.df, (const T *)d.psi, (const T *)d.dt, d.adc, first, last,
                           k2pig, inv2pig, invg, (Cx<T> *)d.mxy, (T *)d.mz, (Cx<T> *)d.states,
                           p->ns);
    else
        hipLaunchKernelGGL((k_plain_excitation<T, false>), grid, block, 0, p->stream, p->ns,
                           (const T *)d.x, (const T *)d.y, (const T *)d.z, (const T *)d.rho,
                           (const T *)d.T1, (const T *)d.T2, (const T *)d.dw, (const T *)d.gx,
                           (const T *)d.gy, (const T *)d.gz, (const Cx<T> *)d.b1,
                           (const T *)d.df, (const T *)d.psi, (const T *)d.dt, d.adc, first, last,
                           k2pig, inv2pig, invg, (Cx<T> *)d.mxy, (T *)d.mz, (Cx<T> *)d.states,
                           p->ns);
    return hipGetLastError();
}

// Stage-1 partials go to `partials` (coil_partial_elements(ns, nadc, nc)
// complex values); `timed` records the plan's kernel timers, which create HIP
// events on first use (the allocation-free entry point passes false).
template <typename T>
hipError_t reduce_t(hom_plan *p, const void *states, int64_t lds, int64_t nadc, void *out,
                    int64_t ldo, void *partials, bool timed) {
    const int64_t nc = p->nc;
    if (nadc == 0 || nc == 0) return hipSuccess;
    if (p->ns == 0) {
        hipLaunchKernelGGL((k_fill_zero<T>), dim3(grid_for(nadc * nc, 256)), dim3(256), 0,
                           p->stream, (Cx<T> *)out, nadc, nc, ldo);
        return hipGetLastError();
    }
    const int64_t ntiles = (p->ns + kSpinTile - 1) / kSpinTile;
    const int64_t nch = nadc * nc;
    const int64_t ktiles = (nadc + 31) / 32, ctiles = (nc + 31) / 32;
    {
        ScopedTimer timer(timed ? p : nullptr, HOM_K_COIL_PARTIALS);
        hipLaunchKernelGGL((k_coil_partials<T>), dim3((unsigned)ntiles, (unsigned)(ktiles * ctiles)),
                           dim3(256), 0, p->stream, p->ns, nadc, nc, (const Cx<T> *)states, lds,
                           (const Cx<T> *)p->d.csm, p->csm_ld ? p->csm_ld : p->ns,
                           (Cx<T> *)partials, nch);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    {
        ScopedTimer timer(timed ? p : nullptr, HOM_K_COIL_FINALIZE);
        hipLaunchKernelGGL((k_coil_finalize<T>), dim3((unsigned)nch), dim3(256), 0, p->stream,
                           ntiles, nadc, (const Cx<T> *)partials, (Cx<T> *)out, ldo);
    }
    return hipGetLastError();
}

// Library-owned stage-1 partials of the allocation-free reduction: one copy
// per device, holding the last call's values until the process ends.  Host
// code never passes this name to a kernel (in host code it is a stand-in, not
// a device address); it uses the address that scratch_address() looked up with
// that device current.
static __device__ __align__(256) unsigned char scratch[HOM_REDUCE_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Partial elements of one reduction, for the smaller of the two spin tiles.
int64_t coil_partial_elements(int64_t ns, int64_t n_adc, int64_t nc) {
    const int64_t ntiles = (ns + kSpinTileMin - 1) / kSpinTileMin;
    return std::max<int64_t>(ntiles, 1) * std::max<int64_t>(n_adc, 1) * std::max<int64_t>(nc, 1);
}

// Partial elements of one fused block (rows of 256 / parts spins).
int64_t fused_partial_elements(int64_t ns, int64_t n_adc, int64_t nc) {
    const int64_t rows = (ns + kFusedSpins - 1) / kFusedSpins * fused_spin_parts(nc);
    return std::max<int64_t>(rows, 1) * std::max<int64_t>(n_adc, 1) * std::max<int64_t>(nc, 1);
}

hipError_t launch_fused_precession(hom_plan *p, int64_t first, int64_t last, int flags,
                                   int64_t n_adc, void *out, int64_t out_ld) {
    return p->is_f32() ? fused_t<float>(p, first, last, flags, n_adc, out, out_ld)
                       : fused_t<double>(p, first, last, flags, n_adc, out, out_ld);
}

hipError_t launch_plain_precession(hom_plan *p, int64_t first, int64_t last, int flags) {
    return p->is_f32() ? precession_t<float>(p, first, last, flags)
                       : precession_t<double>(p, first, last, flags);
}

hipError_t launch_plain_excitation(hom_plan *p, int64_t first, int64_t last, int flags) {
    return p->is_f32() ? excitation_t<float>(p, first, last, flags)
                       : excitation_t<double>(p, first, last, flags);
}

hipError_t launch_coil_reduce(hom_plan *p, const void *states, int64_t ld_states, int64_t n_adc,
                              void *out, int64_t out_ld) {
    return p->is_f32() ? reduce_t<float>(p, states, ld_states, n_adc, out, out_ld, p->d.partials, true)
                       : reduce_t<double>(p, states, ld_states, n_adc, out, out_ld, p->d.partials, true);
}

hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    // the caller made `device` current: the address is that device's copy
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void *scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

hipError_t launch_coil_reduce_scratch(hom_plan *p, const void *states, int64_t ld_states,
                                      int64_t n_adc, void *out, int64_t out_ld) {
    void *ptr = scratch_address(p->device);
    if (!ptr) return hipErrorNotInitialized;
    return p->is_f32() ? reduce_t<float>(p, states, ld_states, n_adc, out, out_ld, ptr, false)
                       : reduce_t<double>(p, states, ld_states, n_adc, out, out_ld, ptr, false);
}

hipError_t launch_fill_complex(hom_plan *p, void *out, int64_t rows, int64_t cols, int64_t ld) {
    if (rows * cols == 0) return hipSuccess;
    if (p->is_f32())
        hipLaunchKernelGGL((k_fill_zero<float>), dim3(grid_for(rows * cols, 256)), dim3(256), 0,
                           p->stream, (Cx<float> *)out, rows, cols, ld);
    else
        hipLaunchKernelGGL((k_fill_zero<double>), dim3(grid_for(rows * cols, 256)), dim3(256), 0,
                           p->stream, (Cx<double> *)out, rows, cols, ld);
    return hipGetLastError();
}

// ---- input statistics of the uploads ----
// Bit patterns of nonnegative doubles order like their values, so an unsigned
// 64-bit atomicMax keeps the largest value.
__device__ inline void atomic_max_nonneg(unsigned long long *a, double v) {
    atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// out[j] (j = 1..16): max over spins of |basis term j|, evaluated as on the
// host (no contraction), so the mixed path picks the same operand exponents.
template <typename T>
__global__ __launch_bounds__(256) void k_basis_absmax(int64_t ns, const T *__restrict__ x, const T *__restrict__ y,
                               const T *__restrict__ z, const T *__restrict__ dw,
                               unsigned long long *__restrict__ out) {
#pragma clang fp contract(off)
    double m[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) m[j] = 0.0;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ns;
         s += (int64_t)gridDim.x * blockDim.x) {
        const double xs = x[s], ys = y[s], zs = z[s], ws = dw[s];
#pragma unroll
        for (int j = 0; j < 16; ++j) m[j] = fmax(m[j], fabs(basis_term(j + 1, xs, ys, zs, ws)));
    }
    // waves of the block, then one atomic per term and block (atomics on the
    // same 16 words serialize, so they are kept few)
    __shared__ double part[4][16];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const double v = wave_max(m[j]);
        if ((threadIdx.x & 63) == 0) part[w][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        const int j = threadIdx.x;
        const double v = fmax(fmax(part[0][j], part[1][j]), fmax(part[2][j], part[3][j]));
        atomic_max_nonneg(out + j + 1, v);
    }
}

// out[c]: max over spins of |csm[s, c]|^2 (column-major, column stride ns).
template <typename T>
__global__ void k_csm_colmax2(int64_t ns, const Cx<T> *__restrict__ csm, unsigned long long *__restrict__ out) {
    const int64_t c = blockIdx.y;
    const Cx<T> *col = csm + c * ns;
    double m = 0.0;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ns;
         s += (int64_t)gridDim.x * blockDim.x) {
        const double re = col[s].re, im = col[s].im;
        m = fmax(m, re * re + im * im);
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(out + c, m);
}

hipError_t launch_basis_absmax(hom_plan *p, unsigned long long *out) {
    const unsigned grid = (unsigned)std::min<int64_t>(grid_for(p->ns, 256), 256);
    if (p->is_f32())
        hipLaunchKernelGGL((k_basis_absmax<float>), dim3(grid), dim3(256), 0, p->stream, p->ns,
                           (const float *)p->d.x, (const float *)p->d.y, (const float *)p->d.z,
                           (const float *)p->d.dw, out);
    else
        hipLaunchKernelGGL((k_basis_absmax<double>), dim3(grid), dim3(256), 0, p->stream, p->ns,
                           (const double *)p->d.x, (const double *)p->d.y, (const double *)p->d.z,
                           (const double *)p->d.dw, out);
    return hipGetLastError();
}

hipError_t launch_csm_colmax2(hom_plan *p, unsigned long long *out) {
    const dim3 grid((unsigned)std::min<int64_t>(grid_for(p->ns, 256), 64), (unsigned)p->nc);
    if (p->is_f32())
        hipLaunchKernelGGL((k_csm_colmax2<float>), grid, dim3(256), 0, p->stream, p->ns,
                           (const Cx<float> *)p->d.csm, out);
    else
        hipLaunchKernelGGL((k_csm_colmax2<double>), grid, dim3(256), 0, p->stream, p->ns,
                           (const Cx<double> *)p->d.csm, out);
    return hipGetLastError();
}

hipError_t launch_reset_state(hom_plan *p) {
    if (p->ns == 0) return hipSuccess;
    if (p->is_f32())
        hipLaunchKernelGGL((k_reset_state<float>), dim3(grid_for(p->ns, 256)), dim3(256), 0,
                           p->stream, p->ns, (const float *)p->d.rho, (Cx<float> *)p->d.mxy,
                           (float *)p->d.mz);
    else
        hipLaunchKernelGGL((k_reset_state<double>), dim3(grid_for(p->ns, 256)), dim3(256), 0,
                           p->stream, p->ns, (const double *)p->d.rho, (Cx<double> *)p->d.mxy,
                           (double *)p->d.mz);
    return hipGetLastError();
}

}  // namespace hom
