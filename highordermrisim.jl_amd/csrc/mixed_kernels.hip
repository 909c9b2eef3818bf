// mixed_kernels.hip -- segment-fused mixed-precision BlochHighOrder path.
//
// Math (restating src/simulation/Bloch/BlochHighOrderSimulatitonMethod.jl:43-94
// and src/simulation/kernels/HighOrderGPU.jl:95-122 in integrated form):
// the trapezoid rule is linear, so the phase of spin s at node n of a
// free-precession segment starting at node n0 is
//     phi(s, n) / 2pi = sum_j b_j(s) * Kr_j[n] + Kr_0[n]
//     Kr_j[n] = -gamma * sum_{m=n0}^{n-1} (h_j[m] + h_j[m+1]) dt_m / 2
// with b_j the masked spatial basis (nominal gradients in place of h1..h3 when
// order one is off) and an extra term b = dw, Kr = -t/2pi for off-resonance.
// Consecutive precession blocks therefore fuse into one segment; only RF
// blocks are sequential barriers.
//
// Cached per (phantom, CSM, mask):
//   k_mix_basis    FP64 basis values in the FP64 MFMA A-operand lane order
//   k_mix_cdec     -log2(e)/T2 per spin, in MFMA output-row order
//   k_mix_bop      B operand of the coil GEMM: split-FP16 (hi, lo) of
//                  csm / S_c (S_c a power-of-two column scale) per 16-spin
//                  group, in MFMA lane order
// Per segment:
//   k_mix_m0max    max |M0| over spins (power-of-two state scale S)
//   k_mix_seg      per spin: phase of M0 in revolutions (FP64) and
//                  log2(|M0| / S) + 14, both in MFMA output-row order
//   k_mix_main     per (16 spins x 16 samples) tile: FP64 phase on the FP64
//                  matrix pipe (v_mfma_f64_16x16x4, accumulator seeded with the
//                  phase of M0), fract/sin/cos/exp2 on VALU, then the complex
//                  coil GEMM as three split-FP16 products (hi*hi + hi*lo +
//                  lo*hi) on v_mfma_f32_16x16x32_f16 with FP32 accumulation
//                  over a spin tile
//   k_mix_finalize FP64 sum of the spin-tile partials in tile order, scales
//                  S * S_c and the spin-independent zeroth-order phase applied
//                  in FP64, written into the signal rows
//   k_mix_end      FP64 end-of-segment state: Mxy <- M0 e^{-T/T2} e^{i phi_end},
//                  Mz <- Mz e^{-T/T1} + rho (1 - e^{-T/T1})
// RF blocks run the f64 per-block excitation kernel unchanged.
#include "hom_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace hom {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int kTermDW = 16;
// Both GEMM operands are pre-scaled by 2^kOpScale (|values| <= 2^14 < FP16 max)
// so that the low halves of the FP16 splits stay in the normal range: FP16
// subnormal inputs may be flushed on the way into the matrix cores, which would
// drop the low half of every element below 1/8.  The finalize removes 2^-2k.
constexpr int kOpScale = 14;
constexpr int kWaves = 4;     // waves per workgroup (each owns kTiles sample tiles)
constexpr int kTiles = 4;     // 16-sample tiles per wave

// Active spatial terms of a field-order mask, in packing order.
struct TermList {
    int n;
    int code[16];  // 1..15: basis index j; kTermDW: off-resonance
};

struct Segment {
    bool rf = false;
    int64_t first = 0, last = 0;   // node range
    int64_t nadc = 0, row0 = 0;    // ADC samples, first signal row
    int64_t tiles = 0;             // 16-sample tiles
    int64_t kr_off = 0, s_off = 0, kend_off = 0;
    double T = 0.0;                // duration
};

struct MixedWorkspace {
    // cache keys
    int basis_flags = -1;
    uint64_t basis_version = 0;
    int table_flags = -1;
    std::vector<int64_t> table_parts;
    uint64_t table_version = 0;

    TermList terms{};
    int NI = 0, NCB = 0;
    int64_t ngroups = 0;
    double *basisA = nullptr;  // ngroups x NI x 64
    float *cdec = nullptr;     // ngroups x 16, permuted
    half8_t *bop = nullptr;    // ngroups x NCB x 2 x 64
    double *csc = nullptr;     // 8 * NCB power-of-two column scales
    double *th0 = nullptr;     // ngroups x 16: phase of M0 (revolutions), permuted
    float *lm = nullptr;       // ngroups x 16: log2(|M0| / S), permuted
    float *m0max = nullptr;    // max |M0| of the current segment
    double *kr = nullptr;      // lane-ordered integrated coefficients
    float *tk = nullptr;       // per-sample time within segment
    double *kr0 = nullptr;     // per-sample zeroth-order phase (revolutions)
    double *kend = nullptr;    // per segment: NT terms, then kr0_end
    float *partials = nullptr;
    int64_t partial_cap = 0;
    std::vector<Segment> segs;
    int gpt = 256;  // 16-spin groups per spin tile
};

namespace {

__device__ __host__ inline double basis_term(int code, double x, double y, double z, double dw) {
    const double x2 = x * x, y2 = y * y, z2 = z * z, r2 = x2 + y2 + z2;
    switch (code) {
        case 1: return x;
        case 2: return y;
        case 3: return z;
        case 4: return x * y;
        case 5: return z * y;
        case 6: return 3.0 * z2 - r2;
        case 7: return x * z;
        case 8: return x2 - y2;
        case 9: return 3.0 * y * x2 - y * y2;
        case 10: return x * z * y;
        case 11: return (5.0 * z2 - r2) * y;
        case 12: return 5.0 * z * z2 - 3.0 * z * r2;
        case 13: return (5.0 * z2 - r2) * x;
        case 14: return x2 * z - y2 * z;
        case 15: return x * x2 - 3.0 * x * y2;
        case kTermDW: return dw;
        default: return 0.0;
    }
}

TermList make_terms(int flags) {
    TermList t{};
    t.n = 0;
    for (int j = 1; j <= 3; ++j) t.code[t.n++] = j;
    if (flags & HOM_HO2)
        for (int j = 4; j <= 8; ++j) t.code[t.n++] = j;
    if (flags & HOM_HO3)
        for (int j = 9; j <= 15; ++j) t.code[t.n++] = j;
    if (flags & HOM_DW_PRECESSION) t.code[t.n++] = kTermDW;
    for (int i = t.n; i < 16; ++i) t.code[i] = 0;
    return t;
}

// Power-of-two column scale >= m (1 when the column is all zero).
__device__ __host__ inline double col_scale(float m) {
    if (!(m > 0.0f)) return 1.0;
    int e = 0;
    frexp((double)m, &e);
    return ldexp(1.0, e);
}

__global__ void k_mix_basis(int64_t ns, int64_t ngroups, int NI, TermList tl,
                            const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, const double *__restrict__ dw,
                            double *__restrict__ basisA) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ngroups * NI * 64) return;
    const int lane = (int)(idx % 64);
    const int64_t gi = idx / 64;
    const int i = (int)(gi % NI);
    const int64_t grp = gi / NI;
    const int64_t s = grp * 16 + lane % 16;   // A row = spin
    const int term = 4 * i + lane / 16;       // A column = term
    double v = 0.0;
    if (s < ns && term < tl.n) v = basis_term(tl.code[term], x[s], y[s], z[s], dw[s]);
    basisA[idx] = v;
}

// cdec[grp*16 + 4g + v] = -log2(e) / T2 of spin grp*16 + g + 4v, the layout in
// which lane l (g = l/16) of the FP64 MFMA output holds its four spin rows.
__global__ void k_mix_cdec(int64_t ns, int64_t ngroups, const double *__restrict__ T2,
                           float *__restrict__ cdec) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ngroups * 16) return;
    const int64_t grp = idx / 16;
    const int r = (int)(idx % 16);
    const int g = r / 4, v = r % 4;
    const int64_t s = grp * 16 + g + 4 * v;
    cdec[idx] = s < ns ? (float)(-1.4426950408889634 / T2[s]) : 0.0f;
}

__global__ void k_fill_f32(float *a, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Makes v an opaque, already rounded register value: the compiler may not
// fold the arithmetic that produced it into later conversions.  Without this
// the backend evaluates (half)(a*b) once as round16(round32(a*b)) and once as
// round16(a*b) (v_fma_mix), and the two can differ by one FP16 ulp.
__device__ inline float opaque(float v) {
    asm volatile("" : "+v"(v));
    return v;
}

// x ~= hi + lo with hi = fp16(x) and lo = fp16(x - hi), both from one value.
__device__ inline void split16(double v, _Float16 &hi, _Float16 &lo) {
    const float vf = opaque((float)v);
    hi = (_Float16)vf;
    lo = (_Float16)(float)(v - (double)hi);
}

// B operand of the coil GEMM for one (group, 16-column block, lane): rows
// k' = 8g + 2v + comp (g = lane/16) are spin g + 4v of the group, real (comp 0)
// or imaginary (comp 1) part of the state; column n = 16 cb + lane%16 is coil
// n/2, real (even n) or imaginary (odd n) part of the signal.
__global__ void k_mix_bop(int64_t ns, int64_t nc, int64_t ngroups, int NCB,
                          const double2 *__restrict__ csm, const double *__restrict__ csc,
                          half8_t *__restrict__ bop) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ngroups * NCB * 64) return;
    const int lane = (int)(idx % 64);
    const int64_t t = idx / 64;
    const int cb = (int)(t % NCB);
    const int64_t grp = t / NCB;
    const int n = 16 * cb + lane % 16;
    const int64_t coil = n / 2;
    const int part = n % 2;
    const int g = lane / 16;
    const double inv = coil < nc ? ldexp(1.0, kOpScale) / csc[coil] : 0.0;
    half8_t hi, lo;
    for (int v = 0; v < 4; ++v) {
        const int64_t s = grp * 16 + g + 4 * v;
        double cre = 0.0, cim = 0.0;
        if (s < ns && coil < nc) {
            const double2 w = csm[s + coil * ns];
            cre = w.x * inv;
            cim = w.y * inv;
        }
        // state re multiplies (re col: cre, im col: cim); state im: (-cim, cre)
        const double b0 = part == 0 ? cre : cim;
        const double b1 = part == 0 ? -cim : cre;
        _Float16 h, l;
        split16(b0, h, l);
        hi[2 * v] = h;
        lo[2 * v] = l;
        split16(b1, h, l);
        hi[2 * v + 1] = h;
        lo[2 * v + 1] = l;
    }
    bop[((grp * NCB + cb) * 2 + 0) * 64 + lane] = hi;
    bop[((grp * NCB + cb) * 2 + 1) * 64 + lane] = lo;
}

// Block-wide max of |M0|, one atomic per workgroup (nonnegative floats order
// like their bit patterns).
__global__ __launch_bounds__(256) void k_mix_m0max(int64_t ns, const double2 *__restrict__ mxy,
                                                   float *__restrict__ out) {
    __shared__ float red[4];
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float v = 0.0f;
    if (s < ns) {
        const double2 m = mxy[s];
        v = (float)sqrt(m.x * m.x + m.y * m.y);
    }
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        if (m > 0.0f) atomicMax(reinterpret_cast<unsigned int *>(out), __float_as_uint(m));
    }
}

// Per spin, in MFMA output-row order (index grp*16 + 4g + v <-> spin grp*16 + g + 4v):
// th0 = arg(M0) / 2pi and lm = log2(|M0| / S) + kOpScale.
__global__ void k_mix_seg(int64_t ns, int64_t ngroups, const double2 *__restrict__ mxy,
                          const float *__restrict__ m0max, double *__restrict__ th0,
                          float *__restrict__ lm) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ngroups * 16) return;
    const int64_t grp = idx / 16;
    const int r = (int)(idx % 16);
    const int64_t s = grp * 16 + r / 4 + 4 * (r % 4);
    double a = 0.0;
    float l = -INFINITY;
    if (s < ns) {
        const double2 m = mxy[s];
        const double mag = sqrt(m.x * m.x + m.y * m.y);
        a = atan2(m.y, m.x) / kTwoPi;
        l = mag > 0.0 ? (float)(log2(mag / col_scale(*m0max)) + kOpScale) : -INFINITY;
    }
    th0[idx] = a;
    lm[idx] = l;
}

// Fused phase + state + coil GEMM.  Workgroup = kWaves waves; wave w owns
// kTiles consecutive 16-sample tiles and loops over the 16-spin groups of one
// spin tile, accumulating a 16-sample x (16 NCB)-column FP32 tile per sample
// tile.  Flat block ids are remapped so that the 8 XCDs each receive a
// contiguous run of work: consecutive work items share a spin tile, so its
// operands are reused from that XCD's L2.
template <int NI, int NCB>
__global__ __launch_bounds__(256) void k_mix_main(
    int64_t ngroups, int gpt, int64_t ntiles, int64_t n_sblocks, int64_t n_work, int cb0,
    int ncb_total, const double *__restrict__ basisA, const float *__restrict__ cdec,
    const half8_t *__restrict__ bop, const double *__restrict__ th0, const float *__restrict__ lm,
    const double *__restrict__ kr, const float *__restrict__ tk, float *__restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int64_t nb = gridDim.x;
    const int64_t bid = blockIdx.x;
    // XCD-aware remap: hardware places block b on XCD b % 8.
    const int64_t per = nb / 8, rem = nb % 8;
    const int64_t xcd = bid % 8, local = bid / 8;
    const int64_t work = xcd * per + std::min<int64_t>(xcd, rem) + local;
    if (work >= n_work) return;
    const int64_t sblock = work % n_sblocks, stile = work / n_sblocks;
    const int64_t tile0 = (sblock * kWaves + w) * kTiles;
    if (tile0 >= ntiles) return;  // no barriers below: idle waves may leave

    double krv[kTiles][NI];
    float tkv[kTiles];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        const int64_t tile = tile0 + t;
        const bool ok = tile < ntiles;
#pragma unroll
        for (int i = 0; i < NI; ++i) krv[t][i] = ok ? kr[(tile * NI + i) * 64 + lane] : 0.0;
        tkv[t] = ok ? tk[tile * 16 + (lane & 15)] : 0.0f;
    }
    floatx4 acc[kTiles][NCB];
#pragma unroll
    for (int t = 0; t < kTiles; ++t)
#pragma unroll
        for (int c = 0; c < NCB; ++c) acc[t][c] = floatx4{0.f, 0.f, 0.f, 0.f};

    const int64_t g0 = stile * gpt;
    const int64_t g1 = std::min<int64_t>(g0 + gpt, ngroups);
    for (int64_t grp = g0; grp < g1; ++grp) {
        double ba[NI];
#pragma unroll
        for (int i = 0; i < NI; ++i) ba[i] = basisA[(grp * NI + i) * 64 + lane];
        const floatx4 cd = *reinterpret_cast<const floatx4 *>(cdec + grp * 16 + 4 * (lane >> 4));
        const floatx4 lmv = *reinterpret_cast<const floatx4 *>(lm + grp * 16 + 4 * (lane >> 4));
        const doublex4 ph0 = *reinterpret_cast<const doublex4 *>(th0 + grp * 16 + 4 * (lane >> 4));
        half8_t bh[NCB], bl[NCB];
#pragma unroll
        for (int c = 0; c < NCB; ++c) {
            bh[c] = bop[((grp * ncb_total + cb0 + c) * 2 + 0) * 64 + lane];
            bl[c] = bop[((grp * ncb_total + cb0 + c) * 2 + 1) * 64 + lane];
        }
#pragma unroll
        for (int t = 0; t < kTiles; ++t) {
            doublex4 ph = ph0;  // phase of M0 seeds the accumulator
#pragma unroll
            for (int i = 0; i < NI; ++i)
                ph = __builtin_amdgcn_mfma_f64_16x16x4f64(ba[i], krv[t][i], ph, 0, 0, 0);
            half8_t ahi, alo;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double r = ph[v] - floor(ph[v]);  // revolutions in [0, 1)
                const float rf = (float)r;
                const float sn = __builtin_amdgcn_sinf(rf);   // sin(2 pi rf)
                const float cs = __builtin_amdgcn_cosf(rf);
                const float d = __builtin_amdgcn_exp2f(fmaf(cd[v], tkv[t], lmv[v]));
                const float re = opaque(d * cs), im = opaque(d * sn);
                const _Float16 hre = (_Float16)re, him = (_Float16)im;
                ahi[2 * v] = hre;
                ahi[2 * v + 1] = him;
                alo[2 * v] = (_Float16)(re - (float)hre);
                alo[2 * v + 1] = (_Float16)(im - (float)him);
            }
#pragma unroll
            for (int c = 0; c < NCB; ++c) {
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bh[c], acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bl[c], acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, bh[c], acc[t][c], 0, 0, 0);
            }
        }
    }
    const int64_t rows = ntiles * 16, cols = 16 * (int64_t)ncb_total;
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        const int64_t tile = tile0 + t;
        if (tile >= ntiles) continue;
#pragma unroll
        for (int c = 0; c < NCB; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t row = tile * 16 + 4 * (lane >> 4) + v;
                const int64_t col = 16 * (cb0 + c) + (lane & 15);
                partials[(stile * rows + row) * cols + col] = acc[t][c][v];
            }
    }
}

// One lane per (sample, coil), coil fastest so a wave reads contiguous
// partial rows; FP64 sum over spin tiles in tile order.
__global__ void k_mix_finalize(int64_t nadc, int64_t nc, int64_t nstiles, int64_t rows, int64_t cols,
                               const float *__restrict__ partials, const float *__restrict__ m0max,
                               const double *__restrict__ csc,
                               const double *__restrict__ kr0, double2 *__restrict__ out, int64_t ld) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nadc * nc) return;
    const int64_t c = idx % nc, k = idx / nc;
    double re = 0.0, im = 0.0;
    for (int64_t st = 0; st < nstiles; ++st) {
        const float2 p = *reinterpret_cast<const float2 *>(partials + (st * rows + k) * cols + 2 * c);
        re += p.x;
        im += p.y;
    }
    const double sc = ldexp(col_scale(*m0max) * csc[c], -2 * kOpScale);
    re *= sc;
    im *= sc;
    const double r = kr0[k] - floor(kr0[k]);
    double sn, cs;
    sincos(kTwoPi * r, &sn, &cs);
    out[k + c * ld] = double2{re * cs - im * sn, re * sn + im * cs};
}

__global__ void k_mix_end(int64_t ns, TermList tl, const double *__restrict__ x,
                          const double *__restrict__ y, const double *__restrict__ z,
                          const double *__restrict__ dw, const double *__restrict__ rho,
                          const double *__restrict__ T1, const double *__restrict__ T2,
                          const double *__restrict__ kend, double T, double2 *__restrict__ mxy,
                          double *__restrict__ mz) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    double rev = kend[tl.n];
    for (int i = 0; i < tl.n; ++i) rev += basis_term(tl.code[i], x[s], y[s], z[s], dw[s]) * kend[i];
    const double r = rev - floor(rev);
    double sn, cs;
    sincos(kTwoPi * r, &sn, &cs);
    const double e2 = exp(-T / T2[s]);
    const double2 m = mxy[s];
    mxy[s] = double2{(m.x * cs - m.y * sn) * e2, (m.x * sn + m.y * cs) * e2};
    const double e1 = exp(-T / T1[s]);
    mz[s] = mz[s] * e1 + rho[s] * (1.0 - e1);
}

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <int NI>
hipError_t launch_main_ni(int ncb, dim3 grid, hipStream_t st, int64_t ngroups, int gpt,
                          int64_t ntiles, int64_t nsb, int64_t nwork, int cb0, int ncb_total,
                          const double *ba, const float *cd, const half8_t *bop, const double *th0,
                          const float *lm, const double *kr, const float *tk, float *part) {
#define HOM_MAIN(NCBV)                                                                          \
    hipLaunchKernelGGL((k_mix_main<NI, NCBV>), grid, dim3(256), 0, st, ngroups, gpt, ntiles, nsb, \
                       nwork, cb0, ncb_total, ba, cd, bop, th0, lm, kr, tk, part);               \
    break;
    switch (ncb) {
        case 1: HOM_MAIN(1)
        case 2: HOM_MAIN(2)
        case 3: HOM_MAIN(3)
        case 4: HOM_MAIN(4)
        default: return hipErrorInvalidValue;
    }
#undef HOM_MAIN
    return hipGetLastError();
}

hipError_t launch_main(int NI, int ncb, dim3 grid, hipStream_t st, int64_t ngroups, int gpt,
                       int64_t ntiles, int64_t nsb, int64_t nwork, int cb0, int ncb_total,
                       const double *ba, const float *cd, const half8_t *bop, const double *th0,
                       const float *lm, const double *kr, const float *tk, float *part) {
    switch (NI) {
        case 1: return launch_main_ni<1>(ncb, grid, st, ngroups, gpt, ntiles, nsb, nwork, cb0, ncb_total, ba, cd, bop, th0, lm, kr, tk, part);
        case 2: return launch_main_ni<2>(ncb, grid, st, ngroups, gpt, ntiles, nsb, nwork, cb0, ncb_total, ba, cd, bop, th0, lm, kr, tk, part);
        case 3: return launch_main_ni<3>(ncb, grid, st, ngroups, gpt, ntiles, nsb, nwork, cb0, ncb_total, ba, cd, bop, th0, lm, kr, tk, part);
        case 4: return launch_main_ni<4>(ncb, grid, st, ngroups, gpt, ntiles, nsb, nwork, cb0, ncb_total, ba, cd, bop, th0, lm, kr, tk, part);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
int dev_upload(hom_plan *p, T **dst, const std::vector<T> &src) {
    if (*dst) hipFree(*dst);
    *dst = nullptr;
    if (src.empty()) return HOM_OK;
    if (hipMalloc((void **)dst, sizeof(T) * src.size()) != hipSuccess) {
        *dst = nullptr;
        return set_error(p, HOM_ERR_MEMORY, "device allocation failed for mixed-path tables");
    }
    return check_hip(p, hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice),
                     "hipMemcpy");
}

// Host FP64 integration of the coefficient waveforms per segment, written in
// the FP64 MFMA B-operand lane order: for 16-sample tile tau and step i,
// lane l holds term 4i + l/16 at sample 16 tau + l%16.
int build_tables(hom_plan *p, MixedWorkspace &W, const int64_t *pf, const int64_t *pl,
                 const uint8_t *rf, int64_t np, int flags) {
    W.segs.clear();
    const TermList &tl = W.terms;
    const int NI = W.NI;
    const int64_t nn = p->nn;
    const bool ho1 = flags & HOM_HO1, ho0 = flags & HOM_HO0;
    // waveform of each packed term (nullptr for the off-resonance term)
    std::vector<const double *> wave(tl.n, nullptr);
    for (int q = 0; q < tl.n; ++q) {
        const int c = tl.code[q];
        if (c == kTermDW) continue;
        if (c <= 3 && !ho1)
            wave[q] = c == 1 ? p->h_gx.data() : c == 2 ? p->h_gy.data() : p->h_gz.data();
        else
            wave[q] = p->h_coef.data() + (size_t)c * nn;
    }
    const double *h0 = p->h_coef.data();
    const double g = p->gamma;
    std::vector<double> kr, kr0, kend;
    std::vector<float> tk;
    int64_t row = 0;
    int64_t b = 0;
    while (b < np) {
        Segment s;
        if (rf[b]) {
            s.rf = true;
            s.first = pf[b];
            s.last = pl[b];
            int64_t nadc = 0;
            for (int64_t n = s.first + 1; n <= s.last; ++n) nadc += p->h_adc[(size_t)n] ? 1 : 0;
            s.nadc = nadc;
            s.row0 = row;
            row += nadc;
            W.segs.push_back(s);
            ++b;
            continue;
        }
        int64_t e = b;
        while (e + 1 < np && !rf[e + 1]) ++e;
        s.first = pf[b];
        s.last = pl[e];
        b = e + 1;
        int64_t nadc = 0;
        for (int64_t n = s.first + 1; n <= s.last; ++n) nadc += p->h_adc[(size_t)n] ? 1 : 0;
        s.nadc = nadc;
        s.row0 = row;
        row += nadc;
        s.tiles = (nadc + 15) / 16;
        s.kr_off = (int64_t)kr.size();
        s.s_off = (int64_t)tk.size();
        s.kend_off = (int64_t)kend.size();
        kr.resize(kr.size() + (size_t)(s.tiles * NI * 64), 0.0);
        tk.resize(tk.size() + (size_t)(s.tiles * 16), 0.0f);
        kr0.resize(kr0.size() + (size_t)(s.tiles * 16), 0.0);
        std::vector<double> I(tl.n, 0.0);
        double I0 = 0.0, t = 0.0;
        int64_t k = 0;
        for (int64_t n = s.first + 1; n <= s.last; ++n) {
            const double dt = p->h_dt[(size_t)(n - 1)];
            for (int q = 0; q < tl.n; ++q)
                if (wave[q]) I[q] += (wave[q][n - 1] + wave[q][n]) * (dt / 2.0);
            if (ho0) I0 += (h0[n - 1] + h0[n]) * (dt / 2.0);
            t += dt;
            if (p->h_adc[(size_t)n]) {
                const int64_t tau = k / 16, j = k % 16;
                for (int q = 0; q < tl.n; ++q) {
                    const double v = tl.code[q] == kTermDW ? -t / kTwoPi : -g * I[q];
                    kr[(size_t)(s.kr_off + (tau * NI + q / 4) * 64 + 16 * (q % 4) + j)] = v;
                }
                tk[(size_t)(s.s_off + k)] = (float)t;
                kr0[(size_t)(s.s_off + k)] = -g * I0;
                ++k;
            }
        }
        for (int q = 0; q < tl.n; ++q) kend.push_back(tl.code[q] == kTermDW ? -t / kTwoPi : -g * I[q]);
        kend.push_back(-g * I0);
        s.T = t;
        W.segs.push_back(s);
    }
    int rc = dev_upload(p, &W.kr, kr);
    rc = rc ? rc : dev_upload(p, &W.tk, tk);
    rc = rc ? rc : dev_upload(p, &W.kr0, kr0);
    rc = rc ? rc : dev_upload(p, &W.kend, kend);
    return rc;
}

}  // namespace

void mixed_free(hom_plan *p) {
    if (!p || !p->mixed) return;
    MixedWorkspace *W = p->mixed;
    hipStreamSynchronize(p->stream);
    void *all[] = {W->basisA, W->cdec, W->bop, W->csc, W->th0, W->lm, W->m0max,
                   W->kr,     W->tk,   W->kr0, W->kend, W->partials};
    for (void *q : all)
        if (q) hipFree(q);
    delete W;
    p->mixed = nullptr;
}

int mixed_run(hom_plan *p, const int64_t *pf, const int64_t *pl, const uint8_t *rf, int64_t np,
              int flags) {
    if (!p->mixed) p->mixed = new MixedWorkspace();
    MixedWorkspace &W = *p->mixed;
    const int64_t ns = p->ns, nc = p->nc;
    int rc = HOM_OK;
    const int mflags = flags & (HOM_HO0 | HOM_HO1 | HOM_HO2 | HOM_HO3 | HOM_DW_PRECESSION);
    const int NCB = (int)((2 * nc + 15) / 16);
    const int64_t ngroups = (ns + 15) / 16;

    // Spin-side operands that depend only on the phantom, the CSM and the mask.
    if (W.basis_flags != mflags || W.ngroups != ngroups || W.NCB != NCB) {
        hipStreamSynchronize(p->stream);
        W.terms = make_terms(mflags);
        W.NI = (W.terms.n + 3) / 4;
        W.NCB = NCB;
        W.ngroups = ngroups;
        for (void **q : {(void **)&W.basisA, (void **)&W.cdec, (void **)&W.bop, (void **)&W.csc,
                         (void **)&W.th0, (void **)&W.lm, (void **)&W.m0max})
            if (*q) { hipFree(*q); *q = nullptr; }
        if (hipMalloc((void **)&W.basisA, sizeof(double) * (size_t)(ngroups * W.NI * 64)) != hipSuccess ||
            hipMalloc((void **)&W.cdec, sizeof(float) * (size_t)(ngroups * 16)) != hipSuccess ||
            hipMalloc((void **)&W.bop, sizeof(half8_t) * (size_t)(ngroups * NCB * 2 * 64)) != hipSuccess ||
            hipMalloc((void **)&W.csc, sizeof(double) * (size_t)(8 * NCB)) != hipSuccess ||
            hipMalloc((void **)&W.th0, sizeof(double) * (size_t)(ngroups * 16)) != hipSuccess ||
            hipMalloc((void **)&W.lm, sizeof(float) * (size_t)(ngroups * 16)) != hipSuccess ||
            hipMalloc((void **)&W.m0max, sizeof(float)) != hipSuccess) {
            W.basis_flags = -1;
            return set_error(p, HOM_ERR_MEMORY, "device allocation failed for mixed-path operands");
        }
        // Power-of-two column scales of the CSM (from the host copy's column maxima).
        std::vector<double> csc((size_t)(8 * NCB), 1.0);
        for (int64_t c = 0; c < nc; ++c) csc[(size_t)c] = col_scale((float)p->h_csm_colmax[(size_t)c]);
        rc = check_hip(p, hipMemcpy(W.csc, csc.data(), sizeof(double) * csc.size(), hipMemcpyHostToDevice),
                       "hipMemcpy");
        if (rc) return rc;
        hipLaunchKernelGGL(k_mix_bop, dim3(blocks_for(ngroups * NCB * 64, 256)), dim3(256), 0,
                           p->stream, ns, nc, ngroups, NCB, (const double2 *)p->d.csm,
                           (const double *)W.csc, W.bop);
        hipLaunchKernelGGL(k_mix_basis, dim3(blocks_for(ngroups * W.NI * 64, 256)), dim3(256), 0,
                           p->stream, ns, ngroups, W.NI, W.terms, (const double *)p->d.x,
                           (const double *)p->d.y, (const double *)p->d.z, (const double *)p->d.dw,
                           W.basisA);
        hipLaunchKernelGGL(k_mix_cdec, dim3(blocks_for(ngroups * 16, 256)), dim3(256), 0, p->stream,
                           ns, ngroups, (const double *)p->d.T2, W.cdec);
        rc = check_hip(p, hipGetLastError(), "mixed basis");
        if (rc) return rc;
        W.basis_flags = mflags;
        W.table_flags = -1;
    }
    // Sequence tables (cached across runs with the same parts and mask).
    std::vector<int64_t> key;
    key.reserve((size_t)(3 * np));
    for (int64_t b = 0; b < np; ++b) {
        key.push_back(pf[b]);
        key.push_back(pl[b]);
        key.push_back(rf[b]);
    }
    if (W.table_flags != mflags || W.table_parts != key) {
        hipStreamSynchronize(p->stream);
        rc = build_tables(p, W, pf, pl, rf, np, mflags);
        if (rc) return rc;
        W.table_flags = mflags;
        W.table_parts = key;
    }
    // Partial buffer sized for the largest segment.
    const int64_t stile_spins = (int64_t)W.gpt * 16;
    const int64_t nstiles = std::max<int64_t>((ns + stile_spins - 1) / stile_spins, 1);
    int64_t need = 0;
    for (const Segment &s : W.segs)
        if (!s.rf) need = std::max<int64_t>(need, nstiles * s.tiles * 16 * 16 * NCB);
    if (need > W.partial_cap) {
        hipStreamSynchronize(p->stream);
        if (W.partials) hipFree(W.partials);
        W.partials = nullptr;
        W.partial_cap = 0;
        if (hipMalloc((void **)&W.partials, sizeof(float) * (size_t)need) != hipSuccess) {
            W.partials = nullptr;
            return set_error(p, HOM_ERR_MEMORY, "device allocation failed for mixed partials");
        }
        W.partial_cap = need;
    }

    double2 *sig = (double2 *)p->signal;
    const int64_t ld = std::max<int64_t>(p->signal_rows, 1);
    for (const Segment &s : W.segs) {
        if (s.rf) {
            // RF blocks: f64 per-block excitation (+ reduction of RF ADC samples).
            if (s.nadc > p->max_adc)
                return set_error(p, HOM_ERR_DIMENSION, "The block acquires more samples than max_adc.");
            rc = check_hip(p, launch_plain_excitation(p, s.first, s.last, flags), "excitation kernel");
            if (rc) return rc;
            if (s.nadc) {
                const int64_t need_p = coil_partial_elements(ns, s.nadc, nc);
                if (need_p > p->d.partial_capacity) {
                    hipStreamSynchronize(p->stream);
                    if (p->d.partials) hipFree(p->d.partials);
                    p->d.partials = nullptr;
                    p->d.partial_capacity = 0;
                    if (hipMalloc(&p->d.partials, (size_t)need_p * p->complex_size()) != hipSuccess)
                        return set_error(p, HOM_ERR_MEMORY, "device allocation failed for coil partials");
                    p->d.partial_capacity = need_p;
                }
                rc = check_hip(p, launch_coil_reduce(p, p->d.states, ns, s.nadc, sig + s.row0, ld),
                               "coil reduction");
                if (rc) return rc;
            }
            continue;
        }
        if (s.nadc > 0 && ns > 0) {
            {
                ScopedTimer timer(p, HOM_K_MIXED_PREP);
                hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, p->stream, W.m0max, (int64_t)1, 0.0f);
                hipLaunchKernelGGL(k_mix_m0max, dim3(blocks_for(ns, 256)), dim3(256), 0, p->stream, ns,
                                   (const double2 *)p->d.mxy, W.m0max);
                hipLaunchKernelGGL(k_mix_seg, dim3(blocks_for(ngroups * 16, 256)), dim3(256), 0,
                                   p->stream, ns, ngroups, (const double2 *)p->d.mxy,
                                   (const float *)W.m0max, W.th0, W.lm);
            }
            rc = check_hip(p, hipGetLastError(), "mixed prep");
            if (rc) return rc;
            const int64_t nsb = (s.tiles + kWaves * kTiles - 1) / (kWaves * kTiles);
            const int64_t nwork = nsb * nstiles;
            for (int cb0 = 0; cb0 < NCB; cb0 += 4) {
                const int ncb = std::min(4, NCB - cb0);
                ScopedTimer timer(p, HOM_K_MIXED_SEGMENT);
                rc = check_hip(p, launch_main(W.NI, ncb, dim3((unsigned)nwork), p->stream, ngroups, W.gpt,
                                              s.tiles, nsb, nwork, cb0, NCB, W.basisA, W.cdec, W.bop,
                                              W.th0, W.lm, W.kr + s.kr_off, W.tk + s.s_off, W.partials),
                               "mixed main kernel");
                if (rc) return rc;
            }
            {
                ScopedTimer timer(p, HOM_K_MIXED_FINALIZE);
                hipLaunchKernelGGL(k_mix_finalize, dim3(blocks_for(s.nadc * nc, 256)), dim3(256), 0,
                                   p->stream, s.nadc, nc, nstiles, s.tiles * 16, (int64_t)(16 * NCB),
                                   (const float *)W.partials, (const float *)W.m0max,
                                   (const double *)W.csc,
                                   (const double *)(W.kr0 + s.s_off), sig + s.row0, ld);
            }
            rc = check_hip(p, hipGetLastError(), "mixed finalize");
            if (rc) return rc;
        } else if (s.nadc > 0) {
            rc = check_hip(p, launch_fill_complex(p, sig + s.row0, s.nadc, nc, ld), "fill");
            if (rc) return rc;
        }
        if (ns > 0) {
            hipLaunchKernelGGL(k_mix_end, dim3(blocks_for(ns, 256)), dim3(256), 0, p->stream, ns,
                               W.terms, (const double *)p->d.x, (const double *)p->d.y,
                               (const double *)p->d.z, (const double *)p->d.dw,
                               (const double *)p->d.rho, (const double *)p->d.T1,
                               (const double *)p->d.T2, (const double *)(W.kend + s.kend_off), s.T,
                               (double2 *)p->d.mxy, (double *)p->d.mz);
            rc = check_hip(p, hipGetLastError(), "mixed segment end");
            if (rc) return rc;
        }
    }
    return HOM_OK;
}

}  // namespace hom
