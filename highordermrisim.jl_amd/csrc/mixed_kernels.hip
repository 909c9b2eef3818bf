// mixed_kernels.hip -- segment-fused mixed-precision BlochHighOrder path.
//
// Math (restating src/simulation/Bloch/BlochHighOrderSimulatitonMethod.jl:43-94
// and src/simulation/kernels/HighOrderGPU.jl:95-122 in integrated form):
// the trapezoid rule is linear, so the phase of spin s at node n of a
// free-precession segment starting at node n0 is
//     phi(s, n) / 2pi = arg(M0_s)/2pi + sum_j b_j(s) * Kr_j[n] + Kr_0[n]
//     Kr_j[n] = -gamma * sum_{m=n0}^{n-1} (h_j[m] + h_j[m+1]) dt_m / 2
// with b_j the masked spatial basis (nominal gradients in place of h1..h3 when
// order one is off) and an extra term b = dw, Kr = -t/2pi for off-resonance.
// Consecutive precession blocks therefore fuse into one segment; only RF
// blocks are sequential barriers.
//
// Phase arithmetic.  The first-order and off-resonance terms (x, y, z, dw)
// carry almost all of the phase (tens of revolutions) and run in FP64 on the
// FP64 matrix pipe, one v_mfma_f64_16x16x4 per 16 spins x 16 samples.  The up
// to 12 higher-order terms are split at sample blocks: their value at the
// block's first sample (the anchor) is computed per spin in FP64 and seeds the
// FP64 accumulator together with arg(M0); only the small offsets from the
// anchor go through split-FP16 MFMAs (hi*hi + lo*hi + hi*lo stacked along K),
// with FP32 accumulation.  FP64 MFMAs do not co-issue with VALU work on
// gfx950; FP16 MFMAs do, so this moves three quarters of the phase work onto
// the pipe that overlaps.  When a segment's offsets would not fit FP16, the
// segment falls back to the all-FP64 phase kernel (k_mix_main64).
//
// Cached per (phantom, CSM, mask):
//   k_mix_basis64 / k_mix_basis  FP64 basis values in FP64-MFMA A-operand order
//   k_mix_aho      split-FP16 higher-order basis, FP16-MFMA A-operand order
//   k_mix_cdec     -log2(e)/T2 per spin, in MFMA output-row order
//   k_mix_bop      B operand of the coil GEMM: split-FP16 (hi, lo) of
//                  csm / S_c (S_c a power-of-two column scale)
// Per segment:
//   k_mix_m0max    max |M0| over spins (power-of-two state scale S)
//   k_mix_seg      per spin: arg(M0) in revolutions (FP64), log2(|M0|/S) + 14
//   k_mix_seed     per (sample block, spin): arg(M0)/2pi + higher-order anchor
//   k_mix_main     per 16 spins x 16 samples: phase (above), fract, v_sin/v_cos
//                  (revolutions), v_exp for |M0| e^{-t/T2}, FP16 split of the
//                  state, complex coil GEMM as three split-FP16 products on
//                  v_mfma_f32_16x16x32_f16, FP32 accumulation over a spin tile
//   k_mix_finalize FP64 sum of the spin-tile partials in tile order, scales
//                  and the spin-independent zeroth-order phase in FP64
//   k_mix_end      FP64 end-of-segment state: Mxy <- M0 e^{-T/T2} e^{i phi_end},
//                  Mz <- Mz e^{-T/T1} + rho (1 - e^{-T/T1})
// RF blocks run the f64 per-block excitation kernel unchanged.
//
// MFMA lane layouts (measured on gfx950): f16 16x16x32 -- A[l%16][8(l/16)+i],
// B[8(l/16)+i][l%16], D[4(l/16)+v][l%16]; f64 16x16x4 -- A[l%16][l/16],
// B[l/16][l%16], D[l/16 + 4v][l%16].  Lane l therefore holds FP64 phases of
// spins g + 4v (g = l/16) at sample l%16; every spin-side array is stored in
// that order, and FP16 A operands use row m <-> spin m/4 + 4(m%4) so both
// outputs line up in registers.
#include "hom_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace hom {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// Both coil-GEMM operands are pre-scaled by 2^kOpScale (|values| <= 2^14 <
// FP16 max) so the low halves of their FP16 splits stay in the normal range.
constexpr int kOpScale = 14;
// Higher-order phase offsets: A = b_j * 2^a_j with max |A| in (2^9, 2^10],
// B = dKr_j * 2^(kHOOut - a_j); the FP32 product sum is offset * 2^kHOOut.
constexpr int kHOA = 10;
constexpr int kHOOut = 20;
// |B| <= 2^9 with |A| <= 2^10 bounds every offset by 0.5 revolutions, so the
// offset can be added to the reduced FP64 phase in FP32 (error <= 1e-7 rev).
// Segments with larger offsets use the all-FP64 phase kernel.
constexpr float kHOBMax = 512.0f;
// Instruction interleave of the per-group compute: after the tile loop the
// scheduler is asked to alternate one MFMA with HOM_SCHED_VALU VALU
// instructions (0 leaves the default schedule), so VALU work of one tile
// issues while the matrix pipe runs the coil products of another.
#ifndef HOM_SCHED_VALU
#define HOM_SCHED_VALU 0
#endif
#ifndef HOM_TILES
#define HOM_TILES 2
#endif
#ifndef HOM_WAVES_PER_EU
#define HOM_WAVES_PER_EU 3
#endif
constexpr int kWaves = 4;            // waves per workgroup
constexpr int kTiles = HOM_TILES;    // 16-sample tiles per wave
constexpr int kBlockTiles = kWaves * kTiles;  // 16-sample tiles per workgroup (sample block)
// Higher-order anchors every kAnchorTiles tiles (256 samples); a workgroup's
// sample block always lies inside one anchor block.
constexpr int kAnchorTiles = 16;
static_assert(kAnchorTiles % kBlockTiles == 0, "anchor blocks must hold whole sample blocks");

// Active spatial terms of a field-order mask, in packing order.
struct TermList {
    int n;
    int code[16];  // 1..15: basis index j; kTermDW: off-resonance
};

// The same terms split into the FP64 group (x, y, z, dw) and higher orders.
struct PhaseSplit {
    int code64[4];   // 1, 2, 3, kTermDW or 0
    int nho;         // number of higher-order terms (0, 5, 7 or 12)
    int nmh;         // FP16 MFMAs per tile for the offsets: ceil(3 nho / 32)
    int codeho[12];
    int aexp[12];    // A-side power-of-two exponent per higher-order term
};

struct Segment {
    bool rf = false;
    bool fallback = false;         // all-FP64 phase for this segment
    int64_t first = 0, last = 0;   // node range
    int64_t nadc = 0, row0 = 0;    // ADC samples, first signal row
    int64_t tiles = 0, nsb = 0;    // 16-sample tiles, sample blocks
    int64_t nab = 0;               // anchor blocks
    int64_t kr_off = 0, k64_off = 0, bho_off = 0, s_off = 0, kend_off = 0, anc_off = 0;
    double T = 0.0;                // duration
};

struct MixedWorkspace {
    int basis_flags = -1;
    int table_flags = -1;
    std::vector<int64_t> table_parts;

    TermList terms{};
    PhaseSplit split{};
    int NI = 0, NCB = 0;
    int64_t ngroups = 0;
    double *basisA = nullptr;   // ngroups x NI x 64 (all-FP64 fallback)
    double *basis64 = nullptr;  // ngroups x 64
    half8_t *aho = nullptr;     // ngroups x nmh x 64
    float *cdec = nullptr;      // ngroups x 16, MFMA output-row order
    half8_t *bop = nullptr;     // ngroups x NCB x 2 x 64
    double *csc = nullptr;      // 8 * NCB power-of-two column scales
    double *th0 = nullptr;      // ngroups x 16: arg(M0) / 2pi
    float *lm = nullptr;        // ngroups x 16: log2(|M0| / S) + kOpScale
    float *m0max = nullptr;     // max |M0| of the current segment
    double *seed = nullptr;     // max_nsb x ngroups x 16
    int64_t seed_cap = 0;
    double *kr = nullptr;       // fallback FP64 tables (all terms)
    double *k64 = nullptr;      // FP64 tables (x, y, z, dw)
    half8_t *bho = nullptr;     // FP16 offset tables
    double *anc = nullptr;      // per segment, per block: higher-order anchors
    float *tk = nullptr;        // per-sample time within segment
    double *kr0 = nullptr;      // per-sample zeroth-order phase (revolutions)
    double *kend = nullptr;     // per segment: all terms, then kr0_end
    float *partials = nullptr;
    int64_t partial_cap = 0;
    std::vector<Segment> segs;
    int gpt = 256;  // 16-spin groups per spin tile
};

namespace {

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

PhaseSplit make_split(int flags, const double *basis_max) {
    PhaseSplit s{};
    s.code64[0] = 1;
    s.code64[1] = 2;
    s.code64[2] = 3;
    s.code64[3] = (flags & HOM_DW_PRECESSION) ? kTermDW : 0;
    s.nho = 0;
    if (flags & HOM_HO2)
        for (int j = 4; j <= 8; ++j) s.codeho[s.nho++] = j;
    if (flags & HOM_HO3)
        for (int j = 9; j <= 15; ++j) s.codeho[s.nho++] = j;
    s.nmh = (3 * s.nho + 31) / 32;
    for (int q = 0; q < 12; ++q) {
        s.aexp[q] = 0;
        if (q < s.nho && basis_max[s.codeho[q]] > 0.0) {
            int e = 0;
            frexp(basis_max[s.codeho[q]], &e);  // max in [2^(e-1), 2^e)
            s.aexp[q] = kHOA - e;
        }
    }
    return s;
}

// Power-of-two column scale >= m (1 when the column is all zero).
__device__ __host__ inline double col_scale(float m) {
    if (!(m > 0.0f)) return 1.0;
    int e = 0;
    frexp((double)m, &e);
    return ldexp(1.0, e);
}

// Spin held by FP16-MFMA A row m: the row order that makes D[4g+v] line up
// with the FP64 output row g + 4v.
__device__ inline int row_spin(int m) { return (m >> 2) + 4 * (m & 3); }

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

__global__ void k_mix_basis64(int64_t ns, int64_t ngroups, PhaseSplit ps,
                              const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ z, const double *__restrict__ dw,
                              double *__restrict__ basis64) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ngroups * 64) return;
    const int lane = (int)(idx % 64);
    const int64_t s = (idx / 64) * 16 + lane % 16;
    const int code = ps.code64[lane / 16];
    basis64[idx] = (s < ns && code) ? basis_term(code, x[s], y[s], z[s], dw[s]) : 0.0;
}

// FP16 A operand of the offset MFMAs.  K slot q (0 <= q < 3 nho) holds term
// q % nho; slots [0, nho) and [2 nho, 3 nho) use the high half of the scaled
// basis value, slots [nho, 2 nho) the low half (paired with hi / hi / lo
// halves of the offsets on the B side).
__global__ void k_mix_aho(int64_t ns, int64_t ngroups, PhaseSplit ps,
                          const double *__restrict__ x, const double *__restrict__ y,
                          const double *__restrict__ z, const double *__restrict__ dw,
                          half8_t *__restrict__ aho) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ngroups * ps.nmh * 64) return;
    const int lane = (int)(idx % 64);
    const int64_t t = idx / 64;
    const int m = (int)(t % ps.nmh);
    const int64_t grp = t / ps.nmh;
    const int64_t s = grp * 16 + row_spin(lane % 16);
    half8_t out;
    for (int i = 0; i < 8; ++i) {
        const int q = 32 * m + 8 * (lane / 16) + i;
        _Float16 v = (_Float16)0.0f;
        if (s < ns && q < 3 * ps.nho) {
            const int term = q % ps.nho, kind = q / ps.nho;
            const double b = ldexp(basis_term(ps.codeho[term], x[s], y[s], z[s], dw[s]), ps.aexp[term]);
            _Float16 hi, lo;
            split16(b, hi, lo);
            v = kind == 1 ? lo : hi;
        }
        out[i] = v;
    }
    aho[idx] = out;
}

// cdec[grp*16 + 4g + v] = -log2(e) / T2 of spin grp*16 + g + 4v.
__global__ void k_mix_cdec(int64_t ns, int64_t ngroups, const double *__restrict__ T2,
                           float *__restrict__ cdec) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ngroups * 16) return;
    const int64_t grp = idx / 16;
    const int r = (int)(idx % 16);
    const int64_t s = grp * 16 + r / 4 + 4 * (r % 4);
    cdec[idx] = s < ns ? (float)(-1.4426950408889634 / T2[s]) : 0.0f;
}

__global__ void k_fill_f32(float *a, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
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

// seed[(ab * ngroups + grp) * 16 + r] = th0 + sum_q b_q * anc[ab][q] (FP64): the
// FP64-accumulator seed of anchor block ab for the spin of row slot r.  One
// thread per spin slot evaluates its basis once and walks the anchor blocks.
__global__ void k_mix_seed(int64_t ns, int64_t ngroups, int64_t nab, PhaseSplit ps,
                           const double *__restrict__ x, const double *__restrict__ y,
                           const double *__restrict__ z, const double *__restrict__ dw,
                           const double *__restrict__ th0, const double *__restrict__ anc,
                           double *__restrict__ seed) {
    const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ngroups * 16) return;
    const int64_t grp = gi / 16;
    const int r = (int)(gi % 16);
    const int64_t s = grp * 16 + r / 4 + 4 * (r % 4);
    double b[12];
    for (int q = 0; q < 12; ++q)
        b[q] = (s < ns && q < ps.nho) ? basis_term(ps.codeho[q], x[s], y[s], z[s], dw[s]) : 0.0;
    const double t0 = th0[gi];
    for (int64_t ab = 0; ab < nab; ++ab) {
        double v = t0;
        for (int q = 0; q < ps.nho; ++q) v += b[q] * anc[ab * ps.nho + q];
        seed[ab * ngroups * 16 + gi] = v;
    }
}

// State element (d cos, d sin) of one spin-sample, split into FP16 halves:
// hi = fp16(x) as one opaque packed register, lo = fp16(x - hi) from it.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

__device__ inline half2_t opaque2(half2_t v) {
    asm volatile("" : "+v"(v));
    return v;
}

struct StateSplit {
    half2_t hi, lo;  // (re, im)
};

// rf: phase in revolutions, already reduced to about [0, 1) (v_sin/v_cos take
// revolutions and accept |rf| <= 256).
__device__ inline StateSplit state_split(float rf, float cd, float tk, float lmv) {
    const float sn = __builtin_amdgcn_sinf(rf);   // sin(2 pi rf)
    const float cs = __builtin_amdgcn_cosf(rf);
    const float d = __builtin_amdgcn_exp2f(fmaf(cd, tk, lmv));
    const float re = d * cs, im = d * sn;
    StateSplit o;
    o.hi = opaque2(half2_t{(_Float16)re, (_Float16)im});
    o.lo = half2_t{(_Float16)(re - (float)o.hi[0]), (_Float16)(im - (float)o.hi[1])};
    return o;
}

// FP64 phase (revolutions) reduced to [0, 1) and rounded to FP32.
__device__ inline float reduce_rev(double ph) { return (float)__builtin_amdgcn_fract(ph); }

// Work-item decode shared by both main kernels.  Hardware places block b on
// XCD b % 8; the remap gives each XCD a contiguous run of work items, and
// consecutive items share a spin tile, so its operands are reused from that
// XCD's L2.
struct WorkItem {
    int64_t sblock, stile, tile0;
    bool valid;
};

__device__ inline WorkItem decode_work(int64_t n_sblocks, int64_t n_work, int64_t ntiles) {
    const int w = threadIdx.x >> 6;
    const int64_t nb = gridDim.x, bid = blockIdx.x;
    const int64_t per = nb / 8, rem = nb % 8;
    const int64_t xcd = bid % 8, local = bid / 8;
    const int64_t work = xcd * per + (xcd < rem ? xcd : rem) + local;
    WorkItem it;
    it.valid = work < n_work;
    it.sblock = work % n_sblocks;
    it.stile = work / n_sblocks;
    it.tile0 = (it.sblock * kWaves + w) * kTiles;
    it.valid = it.valid && it.tile0 < ntiles;
    return it;
}

template <int NCB>
__device__ inline void write_partials(const floatx4 (&acc)[kTiles][NCB], const WorkItem &it,
                                      int64_t ntiles, int cb0, int ncb_total,
                                      float *__restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const int64_t rows = ntiles * 16, cols = 16 * (int64_t)ncb_total;
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        const int64_t tile = it.tile0 + t;
        if (tile >= ntiles) continue;
#pragma unroll
        for (int c = 0; c < NCB; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t row = tile * 16 + 4 * (lane >> 4) + v;
                const int64_t col = 16 * (cb0 + c) + (lane & 15);
                partials[(it.stile * rows + row) * cols + col] = acc[t][c][v];
            }
    }
}

// Fused phase + state + coil GEMM (split phase: FP64 first order + FP16
// higher-order offsets).  Workgroup = kWaves waves over one spin tile; wave w
// owns kTiles consecutive 16-sample tiles and accumulates a 16-sample x
// (16 NCB)-column FP32 tile per sample tile over every 16-spin group of the
// spin tile.
//
// The spin-group operands (coil B operand, higher-order A operand, FP64
// basis, decay rate, log-magnitude, phase seed) are the same for all waves:
// the workgroup copies each group's record into LDS once, double-buffered,
// while it computes the previous group (global loads into staging registers
// before the compute, LDS stores after it, one barrier per group).  Each wave
// keeps its own per-tile FP16 offset operands in a private LDS slice.
template <int NMH, int NCB>
struct SplitLds {
    static constexpr int kBopChunks = NCB * 2 * 64;    // 16-byte chunks
    static constexpr int kAhoChunks = NMH * 64;
    static constexpr int kRecChunks = kBopChunks + kAhoChunks + 32 + 4 + 4 + 8;
    static constexpr int kRecBytes = kRecChunks * 16;
    static constexpr int kBvBytes = kTiles * (NMH > 0 ? NMH : 1) * 64 * 16;  // per wave
    static constexpr int kBytes = 2 * kRecBytes + kWaves * kBvBytes;
    static constexpr int kStage = (kRecChunks + 255) / 256;  // chunks per thread
};

template <int NMH, int NCB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HOM_WAVES_PER_EU, HOM_WAVES_PER_EU))) void k_mix_main(
    int64_t ngroups, int gpt, int64_t ntiles, int64_t n_sblocks, int64_t n_work, int cb0,
    int ncb_total, const double *__restrict__ basis64, const half8_t *__restrict__ aho,
    const float *__restrict__ cdec, const float *__restrict__ lm, const double *__restrict__ seed,
    const half8_t *__restrict__ bop, const double *__restrict__ k64,
    const half8_t *__restrict__ bho, const float *__restrict__ tk, float *__restrict__ partials) {
    using L = SplitLds<NMH, NCB>;
    __shared__ __attribute__((aligned(32))) unsigned char smem[L::kBytes];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const WorkItem it = decode_work(n_sblocks, n_work, ntiles);
    // Every wave of a workgroup has the same work item; only the tile range
    // differs.  Waves without tiles still take part in the copies and barriers.
    {
        const int64_t nb = gridDim.x, bid = blockIdx.x;
        const int64_t per = nb / 8, rem = nb % 8, xcd = bid % 8;
        if (xcd * per + (xcd < rem ? xcd : rem) + bid / 8 >= n_work) return;  // whole workgroup
    }
    const bool active = it.valid;

    // Per-wave FP16 offset operands for its tiles -> private LDS slice.
    uint4 *bvl = reinterpret_cast<uint4 *>(smem + 2 * L::kRecBytes + w * L::kBvBytes);
    double kv[kTiles];
    float tkv[kTiles];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        const int64_t tile = it.tile0 + t;
        const bool ok = active && tile < ntiles;
        kv[t] = ok ? k64[tile * 64 + lane] : 0.0;
        tkv[t] = ok ? tk[tile * 16 + (lane & 15)] : 0.0f;
#pragma unroll
        for (int m = 0; m < NMH; ++m)
            bvl[(t * NMH + m) * 64 + lane] =
                ok ? reinterpret_cast<const uint4 *>(bho)[(tile * NMH + m) * 64 + lane] : uint4{0, 0, 0, 0};
    }

    // Each thread copies chunks tid, tid + 256, ... of every group record.
    // Chunk c of group grp lives at src[j] + grp * stride[j] (bytes).
    const double *seed_b = seed + (it.sblock * kBlockTiles / kAnchorTiles) * ngroups * 16;
    const unsigned char *src[L::kStage];
    int64_t stride[L::kStage];
#pragma unroll
    for (int j = 0; j < L::kStage; ++j) {
        int c = tid + 256 * j;
        const unsigned char *b = reinterpret_cast<const unsigned char *>(bop);
        int64_t st = 0, off = 0;
        if (c >= L::kRecChunks) {
            c = 0;  // unused slot: never loaded
        }
        if (c < L::kBopChunks) {
            b = reinterpret_cast<const unsigned char *>(bop);
            off = ((int64_t)cb0 * 2 * 64 + c) * 16;
            st = (int64_t)ncb_total * 2 * 64 * 16;
        } else if ((c -= L::kBopChunks) < L::kAhoChunks) {
            b = reinterpret_cast<const unsigned char *>(aho);
            off = (int64_t)c * 16;
            st = (int64_t)NMH * 64 * 16;
        } else if ((c -= L::kAhoChunks) < 32) {
            b = reinterpret_cast<const unsigned char *>(basis64);
            off = (int64_t)c * 16;
            st = 64 * 8;
        } else if ((c -= 32) < 4) {
            b = reinterpret_cast<const unsigned char *>(cdec);
            off = (int64_t)c * 16;
            st = 16 * 4;
        } else if ((c -= 4) < 4) {
            b = reinterpret_cast<const unsigned char *>(lm);
            off = (int64_t)c * 16;
            st = 16 * 4;
        } else {
            c -= 4;
            b = reinterpret_cast<const unsigned char *>(seed_b);
            off = (int64_t)c * 16;
            st = 16 * 8;
        }
        src[j] = b + off;
        stride[j] = st;
    }
    uint4 stage[L::kStage];
    auto load_stage = [&](int64_t grp) {
#pragma unroll
        for (int j = 0; j < L::kStage; ++j)
            if (tid + 256 * j < L::kRecChunks)
                stage[j] = *reinterpret_cast<const uint4 *>(src[j] + grp * stride[j]);
    };
    auto store_stage = [&](int buf) {
        uint4 *dst = reinterpret_cast<uint4 *>(smem + buf * L::kRecBytes);
#pragma unroll
        for (int j = 0; j < L::kStage; ++j) {
            const int c = tid + 256 * j;
            if (c < L::kRecChunks) dst[c] = stage[j];
        }
    };

    floatx4 acc[kTiles][NCB];
#pragma unroll
    for (int t = 0; t < kTiles; ++t)
#pragma unroll
        for (int c = 0; c < NCB; ++c) acc[t][c] = floatx4{0.f, 0.f, 0.f, 0.f};

    const int64_t g0 = it.stile * gpt;
    const int64_t g1 = std::min<int64_t>(g0 + gpt, ngroups);
    load_stage(g0);
    store_stage(0);
    __syncthreads();
    for (int64_t grp = g0; grp < g1; ++grp) {
        const int cur = (int)((grp - g0) & 1);
        const bool more = grp + 1 < g1;
        if (more) load_stage(grp + 1);
        if (active) {
            const unsigned char *rec = smem + cur * L::kRecBytes;
            const half8_t *rb = reinterpret_cast<const half8_t *>(rec);
            half8_t bh[NCB], bl[NCB];
#pragma unroll
            for (int c = 0; c < NCB; ++c) {
                bh[c] = rb[(c * 2 + 0) * 64 + lane];
                bl[c] = rb[(c * 2 + 1) * 64 + lane];
            }
            half8_t ah[NMH > 0 ? NMH : 1];
#pragma unroll
            for (int m = 0; m < NMH; ++m) ah[m] = rb[L::kBopChunks + m * 64 + lane];
            const unsigned char *tail = rec + (L::kBopChunks + L::kAhoChunks) * 16;
            const double ba = reinterpret_cast<const double *>(tail)[lane];
            const floatx4 cd = reinterpret_cast<const floatx4 *>(tail + 512)[lane >> 4];
            const floatx4 lmv = reinterpret_cast<const floatx4 *>(tail + 576)[lane >> 4];
            const doublex4 sd = reinterpret_cast<const doublex4 *>(tail + 640)[lane >> 4];
            const half8_t *bv = reinterpret_cast<const half8_t *>(bvl);
#pragma unroll
            for (int t = 0; t < kTiles; ++t) {
                const doublex4 ph = __builtin_amdgcn_mfma_f64_16x16x4f64(ba, kv[t], sd, 0, 0, 0);
                floatx4 off = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int m = 0; m < NMH; ++m)
                    off = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[m], bv[(t * NMH + m) * 64 + lane],
                                                                 off, 0, 0, 0);
                uint4 ahu, alu;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    float rf = reduce_rev(ph[v]);
                    if (NMH > 0) rf = fmaf(off[v], 0x1p-20f, rf);
                    const StateSplit e = state_split(rf, cd[v], tkv[t], lmv[v]);
                    (&ahu.x)[v] = __builtin_bit_cast(unsigned int, e.hi);
                    (&alu.x)[v] = __builtin_bit_cast(unsigned int, e.lo);
                }
                const half8_t ahi = __builtin_bit_cast(half8_t, ahu);
                const half8_t alo = __builtin_bit_cast(half8_t, alu);
#pragma unroll
                for (int c = 0; c < NCB; ++c) {
                    acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bh[c], acc[t][c], 0, 0, 0);
                    acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bl[c], acc[t][c], 0, 0, 0);
                    acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, bh[c], acc[t][c], 0, 0, 0);
                }
            }
#if HOM_SCHED_VALU > 0
#pragma unroll
            for (int i = 0; i < kTiles * (1 + NMH + 3 * NCB); ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);               // one MFMA
                __builtin_amdgcn_sched_group_barrier(0x002, HOM_SCHED_VALU, 0);  // then VALU
            }
#endif
        }
        if (more) store_stage(cur ^ 1);
        __syncthreads();
    }
    if (active) write_partials<NCB>(acc, it, ntiles, cb0, ncb_total, partials);
}

// All-FP64 phase variant (NI FP64 MFMAs over every term, accumulator seeded
// with arg(M0)), used when a segment's higher-order offsets do not fit FP16.
template <int NI, int NCB>
__global__ __launch_bounds__(256) void k_mix_main64(
    int64_t ngroups, int gpt, int64_t ntiles, int64_t n_sblocks, int64_t n_work, int cb0,
    int ncb_total, const double *__restrict__ basisA, const float *__restrict__ cdec,
    const half8_t *__restrict__ bop, const double *__restrict__ th0, const float *__restrict__ lm,
    const double *__restrict__ kr, const float *__restrict__ tk, float *__restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const WorkItem it = decode_work(n_sblocks, n_work, ntiles);
    if (!it.valid) return;

    double krv[kTiles][NI];
    float tkv[kTiles];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        const int64_t tile = it.tile0 + t;
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

    const int64_t g0 = it.stile * gpt;
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
            doublex4 ph = ph0;
#pragma unroll
            for (int i = 0; i < NI; ++i)
                ph = __builtin_amdgcn_mfma_f64_16x16x4f64(ba[i], krv[t][i], ph, 0, 0, 0);
            uint4 ahu, alu;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const StateSplit e = state_split(reduce_rev(ph[v]), cd[v], tkv[t], lmv[v]);
                (&ahu.x)[v] = __builtin_bit_cast(unsigned int, e.hi);
                (&alu.x)[v] = __builtin_bit_cast(unsigned int, e.lo);
            }
            const half8_t ahi = __builtin_bit_cast(half8_t, ahu);
            const half8_t alo = __builtin_bit_cast(half8_t, alu);
#pragma unroll
            for (int c = 0; c < NCB; ++c) {
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bh[c], acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bl[c], acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, bh[c], acc[t][c], 0, 0, 0);
            }
        }
    }
    write_partials<NCB>(acc, it, ntiles, cb0, ncb_total, partials);
}

// One lane per (sample, coil), coil fastest so a wave reads contiguous
// partial rows; FP64 sum over spin tiles in tile order.
__global__ void k_mix_finalize(int64_t nadc, int64_t nc, int64_t nstiles, int64_t rows, int64_t cols,
                               const float *__restrict__ partials, const float *__restrict__ m0max,
                               const double *__restrict__ csc, const double *__restrict__ kr0,
                               double2 *__restrict__ out, int64_t ld) {
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

struct MainArgs {
    dim3 grid;
    hipStream_t st;
    int64_t ngroups;
    int gpt;
    int64_t ntiles, nsb, nwork;
    int cb0, ncb_total;
};

template <int NMH>
hipError_t launch_split_nmh(int ncb, const MainArgs &a, const MixedWorkspace &W, const Segment &s,
                            float *part) {
#define HOM_MAIN(NCBV)                                                                             \
    hipLaunchKernelGGL((k_mix_main<NMH, NCBV>), a.grid, dim3(256), 0, a.st, a.ngroups, a.gpt,     \
                       a.ntiles, a.nsb, a.nwork, a.cb0, a.ncb_total, W.basis64, W.aho, W.cdec,    \
                       W.lm, W.seed, W.bop, W.k64 + s.k64_off, W.bho + s.bho_off,                  \
                       W.tk + s.s_off, part);                                                      \
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

template <int NI>
hipError_t launch_fp64_ni(int ncb, const MainArgs &a, const MixedWorkspace &W, const Segment &s,
                          float *part) {
#define HOM_MAIN(NCBV)                                                                             \
    hipLaunchKernelGGL((k_mix_main64<NI, NCBV>), a.grid, dim3(256), 0, a.st, a.ngroups, a.gpt,    \
                       a.ntiles, a.nsb, a.nwork, a.cb0, a.ncb_total, W.basisA, W.cdec, W.bop,     \
                       W.th0, W.lm, W.kr + s.kr_off, W.tk + s.s_off, part);                        \
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

hipError_t launch_main(int ncb, const MainArgs &a, const MixedWorkspace &W, const Segment &s,
                       float *part) {
    if (!s.fallback) {
        switch (W.split.nmh) {
            case 0: return launch_split_nmh<0>(ncb, a, W, s, part);
            case 1: return launch_split_nmh<1>(ncb, a, W, s, part);
            case 2: return launch_split_nmh<2>(ncb, a, W, s, part);
            default: return hipErrorInvalidValue;
        }
    }
    switch (W.NI) {
        case 1: return launch_fp64_ni<1>(ncb, a, W, s, part);
        case 2: return launch_fp64_ni<2>(ncb, a, W, s, part);
        case 3: return launch_fp64_ni<3>(ncb, a, W, s, part);
        case 4: return launch_fp64_ni<4>(ncb, a, W, s, part);
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

void host_split16(double v, _Float16 &hi, _Float16 &lo) {
    hi = (_Float16)(float)v;
    lo = (_Float16)(float)(v - (double)hi);
}

// Host FP64 integration of the coefficient waveforms per segment, written in
// MFMA B-operand lane order:
//   kr  (all-FP64 fallback): tile tau, step i, lane l -> term 4i + l/16 at sample l%16
//   k64 (FP64 first order):  tile tau, lane l -> term l/16 of (x, y, z, dw) at sample l%16
//   bho (FP16 offsets):      tile tau, MFMA m, lane l, element i -> K slot
//                            32m + 8(l/16) + i at sample l%16
//   anc: per sample block, the higher-order Kr at the block's first sample.
int build_tables(hom_plan *p, MixedWorkspace &W, const int64_t *pf, const int64_t *pl,
                 const uint8_t *rf, int64_t np, int flags) {
    W.segs.clear();
    const TermList &tl = W.terms;
    const PhaseSplit &ps = W.split;
    const int NI = W.NI, NMH = ps.nmh, NHO = ps.nho;
    const int64_t nn = p->nn;
    const bool ho1 = flags & HOM_HO1, ho0 = flags & HOM_HO0;
    auto wave_of = [&](int code) -> const double * {
        if (code == kTermDW || code == 0) return nullptr;
        if (code <= 3 && !ho1) return code == 1 ? p->h_gx.data() : code == 2 ? p->h_gy.data() : p->h_gz.data();
        return p->h_coef.data() + (size_t)code * nn;
    };
    std::vector<const double *> wave(tl.n);
    for (int q = 0; q < tl.n; ++q) wave[q] = wave_of(tl.code[q]);
    // positions of the FP64 and higher-order terms inside the full list
    int pos64[4] = {-1, -1, -1, -1}, posho[12];
    for (int q = 0; q < tl.n; ++q) {
        for (int j = 0; j < 4; ++j)
            if (ps.code64[j] && tl.code[q] == ps.code64[j]) pos64[j] = q;
        for (int j = 0; j < NHO; ++j)
            if (tl.code[q] == ps.codeho[j]) posho[j] = q;
    }
    const double *h0 = p->h_coef.data();
    const double g = p->gamma;
    std::vector<double> kr, k64, kr0, kend, anc;
    std::vector<half8_t> bho;
    std::vector<float> tk;
    int64_t row = 0, b = 0;
    while (b < np) {
        Segment s;
        if (rf[b]) {
            s.rf = true;
            s.first = pf[b];
            s.last = pl[b];
            for (int64_t n = s.first + 1; n <= s.last; ++n) s.nadc += p->h_adc[(size_t)n] ? 1 : 0;
            s.row0 = row;
            row += s.nadc;
            W.segs.push_back(s);
            ++b;
            continue;
        }
        int64_t e = b;
        while (e + 1 < np && !rf[e + 1]) ++e;
        s.first = pf[b];
        s.last = pl[e];
        b = e + 1;
        for (int64_t n = s.first + 1; n <= s.last; ++n) s.nadc += p->h_adc[(size_t)n] ? 1 : 0;
        s.row0 = row;
        row += s.nadc;
        s.tiles = (s.nadc + 15) / 16;
        s.nsb = (s.tiles + kBlockTiles - 1) / kBlockTiles;
        s.nab = (s.tiles + kAnchorTiles - 1) / kAnchorTiles;
        s.kr_off = (int64_t)kr.size();
        s.k64_off = (int64_t)k64.size();
        s.bho_off = (int64_t)bho.size();
        s.s_off = (int64_t)tk.size();
        s.kend_off = (int64_t)kend.size();
        s.anc_off = (int64_t)anc.size();
        kr.resize(kr.size() + (size_t)(s.tiles * NI * 64), 0.0);
        k64.resize(k64.size() + (size_t)(s.tiles * 64), 0.0);
        bho.resize(bho.size() + (size_t)(s.tiles * NMH * 64), half8_t{});
        tk.resize(tk.size() + (size_t)(s.tiles * 16), 0.0f);
        kr0.resize(kr0.size() + (size_t)(s.tiles * 16), 0.0);
        anc.resize(anc.size() + (size_t)(s.nab * NHO), 0.0);
        std::vector<double> I(tl.n, 0.0), Kr(tl.n, 0.0);
        double I0 = 0.0, t = 0.0, bmax = 0.0;
        int64_t k = 0;
        for (int64_t n = s.first + 1; n <= s.last; ++n) {
            const double dt = p->h_dt[(size_t)(n - 1)];
            for (int q = 0; q < tl.n; ++q)
                if (wave[q]) I[q] += (wave[q][n - 1] + wave[q][n]) * (dt / 2.0);
            if (ho0) I0 += (h0[n - 1] + h0[n]) * (dt / 2.0);
            t += dt;
            if (!p->h_adc[(size_t)n]) continue;
            for (int q = 0; q < tl.n; ++q) Kr[q] = tl.code[q] == kTermDW ? -t / kTwoPi : -g * I[q];
            const int64_t tau = k / 16, j = k % 16, sb = k / (16 * kAnchorTiles);
            for (int q = 0; q < tl.n; ++q)
                kr[(size_t)(s.kr_off + (tau * NI + q / 4) * 64 + 16 * (q % 4) + j)] = Kr[q];
            for (int q = 0; q < 4; ++q)
                k64[(size_t)(s.k64_off + tau * 64 + 16 * q + j)] = pos64[q] >= 0 ? Kr[pos64[q]] : 0.0;
            if (k % (16 * kAnchorTiles) == 0)
                for (int q = 0; q < NHO; ++q) anc[(size_t)(s.anc_off + sb * NHO + q)] = Kr[posho[q]];
            for (int slot = 0; slot < 3 * NHO; ++slot) {
                const int term = slot % NHO, kind = slot / NHO;
                const double bval = ldexp(Kr[posho[term]] - anc[(size_t)(s.anc_off + sb * NHO + term)],
                                          kHOOut - ps.aexp[term]);
                bmax = std::max(bmax, std::fabs(bval));
                _Float16 hi, lo;
                host_split16(bval, hi, lo);
                const int m = slot / 32, sl = slot % 32;
                bho[(size_t)(s.bho_off + (tau * NMH + m) * 64 + 16 * (sl / 8) + j)][sl % 8] =
                    kind == 2 ? lo : hi;
            }
            tk[(size_t)(s.s_off + k)] = (float)t;
            kr0[(size_t)(s.s_off + k)] = -g * I0;
            ++k;
        }
        s.fallback = !(bmax <= kHOBMax);
        for (int q = 0; q < tl.n; ++q) kend.push_back(tl.code[q] == kTermDW ? -t / kTwoPi : -g * I[q]);
        kend.push_back(-g * I0);
        s.T = t;
        W.segs.push_back(s);
    }
    int rc = dev_upload(p, &W.kr, kr);
    rc = rc ? rc : dev_upload(p, &W.k64, k64);
    rc = rc ? rc : dev_upload(p, &W.bho, bho);
    rc = rc ? rc : dev_upload(p, &W.anc, anc);
    rc = rc ? rc : dev_upload(p, &W.tk, tk);
    rc = rc ? rc : dev_upload(p, &W.kr0, kr0);
    rc = rc ? rc : dev_upload(p, &W.kend, kend);
    return rc;
}

template <typename T>
void free_dev(T *&q) {
    if (q) hipFree(q);
    q = nullptr;
}

}  // namespace

void mixed_free(hom_plan *p) {
    if (!p || !p->mixed) return;
    MixedWorkspace *W = p->mixed;
    hipStreamSynchronize(p->stream);
    free_dev(W->basisA); free_dev(W->basis64); free_dev(W->aho); free_dev(W->cdec);
    free_dev(W->bop); free_dev(W->csc); free_dev(W->th0); free_dev(W->lm); free_dev(W->m0max);
    free_dev(W->seed); free_dev(W->kr); free_dev(W->k64); free_dev(W->bho); free_dev(W->anc);
    free_dev(W->tk); free_dev(W->kr0); free_dev(W->kend); free_dev(W->partials);
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
    const double *dx = (const double *)p->d.x, *dy = (const double *)p->d.y,
                 *dz = (const double *)p->d.z, *ddw = (const double *)p->d.dw;

    // Spin-side operands that depend only on the phantom, the CSM and the mask.
    if (W.basis_flags != mflags || W.ngroups != ngroups || W.NCB != NCB) {
        hipStreamSynchronize(p->stream);
        W.terms = make_terms(mflags);
        W.split = make_split(mflags, p->h_basis_max);
        W.NI = (W.terms.n + 3) / 4;
        W.NCB = NCB;
        W.ngroups = ngroups;
        free_dev(W.basisA); free_dev(W.basis64); free_dev(W.aho); free_dev(W.cdec);
        free_dev(W.bop); free_dev(W.csc); free_dev(W.th0); free_dev(W.lm); free_dev(W.m0max);
        const int nmh = std::max(W.split.nmh, 1);
        if (hipMalloc((void **)&W.basisA, sizeof(double) * (size_t)(ngroups * W.NI * 64)) != hipSuccess ||
            hipMalloc((void **)&W.basis64, sizeof(double) * (size_t)(ngroups * 64)) != hipSuccess ||
            hipMalloc((void **)&W.aho, sizeof(half8_t) * (size_t)(ngroups * nmh * 64)) != hipSuccess ||
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
                           p->stream, ns, ngroups, W.NI, W.terms, dx, dy, dz, ddw, W.basisA);
        hipLaunchKernelGGL(k_mix_basis64, dim3(blocks_for(ngroups * 64, 256)), dim3(256), 0,
                           p->stream, ns, ngroups, W.split, dx, dy, dz, ddw, W.basis64);
        if (W.split.nmh > 0)
            hipLaunchKernelGGL(k_mix_aho, dim3(blocks_for(ngroups * W.split.nmh * 64, 256)), dim3(256),
                               0, p->stream, ns, ngroups, W.split, dx, dy, dz, ddw, W.aho);
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
    // Partial and seed buffers sized for the largest segment.
    const int64_t stile_spins = (int64_t)W.gpt * 16;
    const int64_t nstiles = std::max<int64_t>((ns + stile_spins - 1) / stile_spins, 1);
    int64_t need = 0, need_seed = 0;
    for (const Segment &s : W.segs)
        if (!s.rf) {
            need = std::max<int64_t>(need, nstiles * s.tiles * 16 * 16 * NCB);
            need_seed = std::max<int64_t>(need_seed, s.nab * ngroups * 16);
        }
    if (need > W.partial_cap || need_seed > W.seed_cap) {
        hipStreamSynchronize(p->stream);
        if (need > W.partial_cap) {
            free_dev(W.partials);
            W.partial_cap = 0;
            if (hipMalloc((void **)&W.partials, sizeof(float) * (size_t)need) != hipSuccess) {
                W.partials = nullptr;
                return set_error(p, HOM_ERR_MEMORY, "device allocation failed for mixed partials");
            }
            W.partial_cap = need;
        }
        if (need_seed > W.seed_cap) {
            free_dev(W.seed);
            W.seed_cap = 0;
            if (hipMalloc((void **)&W.seed, sizeof(double) * (size_t)need_seed) != hipSuccess) {
                W.seed = nullptr;
                return set_error(p, HOM_ERR_MEMORY, "device allocation failed for phase seeds");
            }
            W.seed_cap = need_seed;
        }
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
                if (!s.fallback)
                    hipLaunchKernelGGL(k_mix_seed, dim3(blocks_for(ngroups * 16, 256)), dim3(256),
                                       0, p->stream, ns, ngroups, s.nab, W.split, dx, dy, dz, ddw,
                                       (const double *)W.th0, (const double *)(W.anc + s.anc_off),
                                       W.seed);
            }
            rc = check_hip(p, hipGetLastError(), "mixed prep");
            if (rc) return rc;
            const int64_t nwork = s.nsb * nstiles;
            for (int cb0 = 0; cb0 < NCB; cb0 += 4) {
                const int ncb = std::min(4, NCB - cb0);
                MainArgs a{dim3((unsigned)nwork), p->stream, ngroups, W.gpt, s.tiles, s.nsb, nwork,
                           cb0, NCB};
                ScopedTimer timer(p, HOM_K_MIXED_SEGMENT);
                rc = check_hip(p, launch_main(ncb, a, W, s, W.partials), "mixed main kernel");
                if (rc) return rc;
            }
            {
                ScopedTimer timer(p, HOM_K_MIXED_FINALIZE);
                hipLaunchKernelGGL(k_mix_finalize, dim3(blocks_for(s.nadc * nc, 256)), dim3(256), 0,
                                   p->stream, s.nadc, nc, nstiles, s.tiles * 16, (int64_t)(16 * NCB),
                                   (const float *)W.partials, (const float *)W.m0max,
                                   (const double *)W.csc, (const double *)(W.kr0 + s.s_off),
                                   sig + s.row0, ld);
            }
            rc = check_hip(p, hipGetLastError(), "mixed finalize");
            if (rc) return rc;
        } else if (s.nadc > 0) {
            rc = check_hip(p, launch_fill_complex(p, sig + s.row0, s.nadc, nc, ld), "fill");
            if (rc) return rc;
        }
        if (ns > 0) {
            hipLaunchKernelGGL(k_mix_end, dim3(blocks_for(ns, 256)), dim3(256), 0, p->stream, ns,
                               W.terms, dx, dy, dz, ddw, (const double *)p->d.rho,
                               (const double *)p->d.T1, (const double *)p->d.T2,
                               (const double *)(W.kend + s.kend_off), s.T, (double2 *)p->d.mxy,
                               (double *)p->d.mz);
            rc = check_hip(p, hipGetLastError(), "mixed segment end");
            if (rc) return rc;
        }
    }
    return HOM_OK;
}

}  // namespace hom
