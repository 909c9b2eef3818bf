// mixed_common.h -- types, host table builder and device bodies of the
// segment-fused mixed-precision BlochHighOrder path, shared by the per-plan
// launches (mixed_kernels.hip) and the batched serving steps (serve.hip).
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
// Phase arithmetic: FP64 anchors, FP32 offsets.
//   * Higher orders (up to 12 terms) are anchored per spin every 256 samples:
//     the seed (k_mix_seed, FP64) holds arg(M0)/2pi plus their value at the
//     anchor block's first sample.
//   * The first-order and off-resonance terms (x, y, z, dw) carry almost all
//     of the phase (tens of revolutions).  They are anchored every 16
//     samples: at the reference sample 8 of each 16-sample tile the FP64
//     phase seed + sum_j b_j Kr_j[ref] is one v_mfma_f64_16x16x4 per 16 spins
//     x 16 references, i.e. per 16-spin group and workgroup sample block of
//     16 tiles, reduced to [0, 1) revolutions and kept in LDS as FP32.
//   * Everything between anchors is an offset: higher orders from the
//     256-sample anchor, first order from the tile's reference (at most 8
//     samples away).  All offsets of a 16 x 16 tile come out of one or two
//     split-FP16 MFMAs with FP32 accumulation: higher orders as hi*hi + lo*hi
//     + hi*lo, first order as the six products of a three-way split whose
//     index sum is at most 2 (33-bit operands).
//   Per tile the phase is then anchor + offset in FP32.  FP64 MFMAs do not
//   co-issue with VALU work on gfx950 and hold the VALU for their 64 cycles;
//   one per group and sample block instead of one per tile is 1/8 of that.
//   When a segment's offsets would not fit FP16, the segment falls back to
//   the all-FP64 phase body (main64_body).
//
// MFMA lane layouts (measured on gfx950): f16 16x16x32 -- A[l%16][8(l/16)+i],
// B[8(l/16)+i][l%16], D[4(l/16)+v][l%16]; f64 16x16x4 -- A[l%16][l/16],
// B[l/16][l%16], D[l/16 + 4v][l%16].  Lane l of an FP64 output therefore
// holds spins g + 4v (g = l/16) in column l%16; every spin-side array is
// stored in that order, and FP16 A operands use row m <-> spin m/4 + 4(m%4)
// so the FP16 outputs D[4g + v] line up with it.
#pragma once

#include "hom_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace hom {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// Both coil-GEMM operands are pre-scaled by 2^kOpScale (|values| <= 2^14 <
// FP16 max) so the low halves of their FP16 splits stay in the normal range.
constexpr int kOpScale = 14;
// Higher-order phase offsets: A = b_j * 2^a_j with max |A| in (2^9, 2^10],
// B = dKr_j * 2^(kHOOut - a_j); the FP32 product sum is offset * 2^kHOOut.
constexpr int kHOA = 10;
constexpr int kHOOut = 20;
// |B| <= 2^13 with |A| <= 2^10 bounds every higher-order offset by 8
// revolutions within its 256-sample anchor block; segments with larger
// offsets use the all-FP64 phase body.  As for the first order below, the
// bound only selects the body: on the 1M-spin, 32-coil, 16 x 8000-sample
// spiral with its 2nd- and 3rd-order fields scaled 1x / 10x / 30x / 100x the
// split path measured NRMSE 0.88 / 0.94 / 0.92 / 0.97e-6 against the f64
// oracle, where a bound of 2^9 (0.5 revolution) sent 10x and more to the
// all-FP64 body (1.45x the kernel time).  With the first-order bound, a
// phase stays below 12 x 8 + 4 x 16 + 1 = 161 revolutions, inside the
// 256-revolution input range of v_sin_f32 / v_cos_f32.
constexpr float kHOBMax = 8192.0f;
// First-order offsets: A = b_j * 2^a_j with max |A| in (2^13, 2^14], B =
// dKr_j * 2^(kHOOut - a_j) (the same product scale, one accumulator), split
// three ways; |B| <= 2^10 bounds each term by 16 revolutions (a spin 0.1 m
// from the centre turns about 2 revolutions in 8 samples of 2 us at 30 mT/m).
// The bound only selects the body: the split operands stay in FP16's normal
// range up to it, and the FP32 sums of larger offsets round more coarsely --
// measured on the 1M-spin, 32-coil, 16 x 8000-sample spiral with its
// gradients scaled 1x / 2x / 4x / 6x: NRMSE 0.9 / 1.0 / 1.4 / 1.9e-6 against
// the f64 oracle, where a bound of 2^8 sent 2x and more to the all-FP64 body
// (1.45x the kernel time).
constexpr int kFOA = 14;
constexpr float kFOBMax = 1024.0f;
// First-order references per 16-sample tile (sample 8).
constexpr int kRefsPerTile = 1;
constexpr int kRefHalf = 16 / kRefsPerTile;
// Four 16-sample tiles per wave at two waves per SIMD (up to 256 VGPRs):
// each group record and coil operand read from LDS then serves 64 samples
// per wave.  Two tiles per wave at three waves per SIMD (168 VGPRs) ran 5 %
// slower, and one tile per wave with eight waves per workgroup (four per
// SIMD) 11 % slower: the group loop is bound by LDS bytes per spin-sample
// more than by occupancy.
#ifndef HOM_TILES
#define HOM_TILES 4
#endif
#ifndef HOM_WAVES_PER_EU
#define HOM_WAVES_PER_EU 2
#endif
constexpr int kWaves = 4;                      // waves per workgroup
constexpr int kTiles = HOM_TILES;              // 16-sample tiles per wave
constexpr int kBlockTiles = kWaves * kTiles;   // 16-sample tiles per workgroup (sample block)
// Higher-order anchors every kAnchorTiles tiles (256 samples); a workgroup's
// sample block always lies inside one anchor block.
constexpr int kAnchorTiles = 16;
constexpr int kBlocksPerAnchor = kAnchorTiles / kBlockTiles;
static_assert(kAnchorTiles % kBlockTiles == 0, "anchor blocks must hold whole sample blocks");
// The first-order references of a sample block are the 16 columns of one
// FP64 MFMA.
static_assert(kBlockTiles * kRefsPerTile == 16, "one FP64 MFMA per group and sample block");

// Sample (within the segment) whose first-order Kr anchors part h of tile
// tau: the middle sample of the part, clamped to the segment's last sample.
inline int64_t ref_sample(int64_t tau, int h, int64_t nadc) {
    const int64_t r = tau * 16 + h * kRefHalf + kRefHalf / 2;
    return r < nadc ? r : nadc - 1;
}

// Active spatial terms of a field-order mask, in packing order.
struct TermList {
    int n;
    int code[16];  // 1..15: basis index j; kTermDW: off-resonance
};

// The same terms split into the first-order group (x, y, z, dw: FP64
// anchors, three-way split offsets) and higher orders (256-sample anchors,
// two-way split offsets).  Offset K slots: [0, 3 nho) higher order (slot
// kind * nho + term, kind 0..2 = A hi/B hi, A lo/B hi, A hi/B lo), then
// [3 nho, 3 nho + 6 n1) first order (slot 3 nho + p * n1 + term, product p
// of kFOParts).
struct PhaseSplit {
    int code64[4];   // 1, 2, 3, kTermDW or 0 (FP64 anchor MFMA: K = 4)
    int n1;          // first-order terms with a nonzero code (3 or 4)
    int code1[4];    // their codes
    int aexp1[4];    // A-side power-of-two exponent per first-order term
    int nho;         // number of higher-order terms (0, 5, 7 or 12)
    int nmh;         // FP16 MFMAs per tile for the offsets: ceil((3 nho + 6 n1) / 32)
    int codeho[12];
    int aexp[12];    // A-side power-of-two exponent per higher-order term
};

// Three-way split products: A part and B part (0 hi, 1 mid, 2 lo) of product
// p; every pair with index sum <= 2.
__host__ __device__ inline int fo_apart(int p) { return p == 2 || p == 4 ? 1 : p == 5 ? 2 : 0; }
__host__ __device__ inline int fo_bpart(int p) { return p == 1 || p == 4 ? 1 : p == 3 ? 2 : 0; }
constexpr int kFOParts = 6;

struct Segment {
    bool rf = false;
    bool fallback = false;         // all-FP64 phase for this segment
    int64_t first = 0, last = 0;   // node range
    int64_t nadc = 0, row0 = 0;    // ADC samples, first signal row
    int64_t tiles = 0, nsb = 0;    // 16-sample tiles, sample blocks
    int64_t nab = 0;               // anchor blocks
    int64_t kr_off = 0, kref_off = 0, bho_off = 0, s_off = 0, kend_off = 0, anc_off = 0;
    double T = 0.0;                // duration
};

// Host view of a sequence's node arrays (coef: 16 x nn, row j contiguous).
struct SeqHost {
    int64_t nn = 0;
    double gamma = 0.0;
    const double *dt = nullptr, *gx = nullptr, *gy = nullptr, *gz = nullptr, *coef = nullptr;
    const uint8_t *adc = nullptr;
};

// Host FP64 tables of one run, concatenated over its segments.
struct TableSet {
    std::vector<double> kr, kref, kr0, kend, anc;
    std::vector<half8_t> bho;
    std::vector<float> tk;
};

TermList make_terms(int flags);
PhaseSplit make_split(int flags, const double *basis_max);
// Segments (consecutive precession parts fused) and their integrated tables,
// in two steps: plan() lists the segments with their sizes and table offsets
// and sizes every table; fill(i) computes precession segment i's entries (and
// decides its phase body), so a caller can fill segment i + 1 on the host
// while the device runs segment i.  The all-FP64 table (kr) grows by each
// falling-back segment as it is filled, within kr_bound() entries.
class TableBuilder {
  public:
    TableBuilder(const SeqHost &sq, const TermList &tl, const PhaseSplit &ps, int NI, int flags,
                 std::vector<Segment> &segs, TableSet &ts)
        : sq_(sq), tl_(tl), ps_(ps), NI_(NI), flags_(flags), segs_(segs), ts_(ts) {}
    void plan(const int64_t *pf, const int64_t *pl, const uint8_t *rf, int64_t np);
    void fill(size_t i);
    int64_t kr_bound() const { return kr_bound_; }

  private:
    const SeqHost &sq_;
    const TermList &tl_;
    const PhaseSplit &ps_;
    int NI_, flags_;
    std::vector<Segment> &segs_;
    TableSet &ts_;
    int64_t kr_bound_ = 0;
};
// Both steps for every segment.
void build_tables(const SeqHost &sq, const TermList &tl, const PhaseSplit &ps, int NI, int flags,
                  const int64_t *pf, const int64_t *pl, const uint8_t *rf, int64_t np,
                  std::vector<Segment> &segs, TableSet &ts);

// ---------------------------------------------------------------- device --

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

// Spin of MFMA output-row slot r (0..15) of a 16-spin group.
__device__ inline int slot_spin(int r) { return r / 4 + 4 * (r % 4); }

// Makes v an opaque, already rounded register value: the compiler may not
// fold the arithmetic that produced it into later conversions.  Without this
// the backend evaluates (half)(a*b) once as round16(round32(a*b)) and once as
// round16(a*b) (v_fma_mix), and the two can differ by one FP16 ulp.
__device__ inline float opaque(float v) {
    asm volatile("" : "+v"(v));
    return v;
}

__device__ inline half2_t opaque2(half2_t v) {
    asm volatile("" : "+v"(v));
    return v;
}

// x ~= hi + lo with hi = fp16(x) and lo = fp16(x - hi), both from one value.
__device__ inline void split16(double v, _Float16 &hi, _Float16 &lo) {
    const float vf = opaque((float)v);
    hi = (_Float16)vf;
    lo = (_Float16)(float)(v - (double)hi);
}

// x ~= p[0] + p[1] + p[2], each part fp16 of what the previous ones leave
// (the differences are exact in FP64).
__device__ inline void split16x3(double v, _Float16 (&p)[3]) {
    p[0] = (_Float16)opaque((float)v);
    const double r1 = v - (double)p[0];
    p[1] = (_Float16)opaque((float)r1);
    p[2] = (_Float16)(float)(r1 - (double)p[1]);
}

// ---- spin-side operands (one element per call; idx ranges in comments) ----

// idx < ngroups * NI * 64: FP64 A operand of all terms (all-FP64 fallback).
__device__ inline void basisA_elem(int64_t idx, int64_t ns, int NI, const TermList &tl,
                                   const double *x, const double *y, const double *z,
                                   const double *dw, double *basisA) {
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

// idx < (ngroups + 1) * 64: FP64 A operand of (x, y, z, dw); the extra group
// is zero padding (a workgroup stages the next group's operand with every
// group record).
__device__ inline void basis64_elem(int64_t idx, int64_t ns, const PhaseSplit &ps, const double *x,
                                    const double *y, const double *z, const double *dw,
                                    double *basis64) {
    const int lane = (int)(idx % 64);
    const int64_t s = (idx / 64) * 16 + lane % 16;
    const int code = ps.code64[lane / 16];
    basis64[idx] = (s < ns && code) ? basis_term(code, x[s], y[s], z[s], dw[s]) : 0.0;
}

// idx < ngroups * nmh * 64: FP16 A operand of the offset MFMAs.  Higher-order
// K slot q (0 <= q < 3 nho) holds term q % nho: slots [0, nho) and
// [2 nho, 3 nho) the high half of the scaled basis value, slots [nho, 2 nho)
// the low half (paired with hi / hi / lo halves of the offsets on the B side).
// First-order slot 3 nho + p n1 + term holds part fo_apart(p) of the
// three-way split of the scaled basis value.
__device__ inline void aho_elem(int64_t idx, int64_t ns, const PhaseSplit &ps, const double *x,
                                const double *y, const double *z, const double *dw, half8_t *aho) {
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
        } else if (s < ns && q < 3 * ps.nho + kFOParts * ps.n1) {
            const int r = q - 3 * ps.nho, term = r % ps.n1, prod = r / ps.n1;
            const double b = ldexp(basis_term(ps.code1[term], x[s], y[s], z[s], dw[s]), ps.aexp1[term]);
            _Float16 parts[3];
            split16x3(b, parts);
            v = parts[fo_apart(prod)];
        }
        out[i] = v;
    }
    aho[idx] = out;
}

// idx < ngroups * 16: cdec[grp*16 + 4g + v] = -log2(e) / T2 of spin grp*16 + g + 4v.
__device__ inline void cdec_elem(int64_t idx, int64_t ns, const double *T2, float *cdec) {
    const int64_t s = (idx / 16) * 16 + slot_spin((int)(idx % 16));
    cdec[idx] = s < ns ? (float)(-1.4426950408889634 / T2[s]) : 0.0f;
}

// idx < ngroups * NCB * 64: B operand of the coil GEMM for one (group,
// 16-column block, lane): rows k' = 8g + 2v + comp (g = lane/16) are spin
// g + 4v of the group, real (comp 0) or imaginary (comp 1) part of the state;
// column n = 16 cb + lane%16 is coil n/2, real (even n) or imaginary (odd n)
// part of the signal.  csm has column stride ldc.
__device__ inline void bop_elem(int64_t idx, int64_t ns, int64_t nc, int NCB, const double2 *csm,
                                int64_t ldc, const double *csc, half8_t *bop) {
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
            const double2 w = csm[s + coil * ldc];
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

// ---- per-segment spin-side steps ----

// Workgroup (256 threads) maximum of v >= 0 into *out by one atomic
// (non-negative floats order like their bit patterns).
__device__ inline void m0max_reduce(float v, float *out) {
    __shared__ float red[4];
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        if (m > 0.0f) atomicMax(reinterpret_cast<unsigned int *>(out), __float_as_uint(m));
    }
    __syncthreads();  // red is reused by the next call of a persistent block
}

// Block-wide max of |M0| over spins s = base + threadIdx.x.
__device__ inline void m0max_block(int64_t s, int64_t ns, const double2 *mxy, float *out) {
    float v = 0.0f;
    if (s < ns) {
        const double2 m = mxy[s];
        v = (float)sqrt(m.x * m.x + m.y * m.y);
    }
    m0max_reduce(v, out);
}

// idx < ngroups * 16, MFMA output-row order: th0 = arg(M0) / 2pi and
// lm = log2(|M0| / S) + kOpScale.
__device__ inline void seg_elem(int64_t idx, int64_t ns, const double2 *mxy, float m0max,
                                double *th0, float *lm) {
    const int64_t s = (idx / 16) * 16 + slot_spin((int)(idx % 16));
    double a = 0.0;
    float l = -INFINITY;
    if (s < ns) {
        const double2 m = mxy[s];
        const double mag = sqrt(m.x * m.x + m.y * m.y);
        a = atan2(m.y, m.x) / kTwoPi;
        l = mag > 0.0 ? (float)(log2(mag / col_scale(m0max)) + kOpScale) : -INFINITY;
    }
    th0[idx] = a;
    lm[idx] = l;
}

// gi < ngroups * 16: seed[(ab * ngroups) * 16 + gi] = th0 + sum_q b_q * anc[ab][q]
// (FP64) for ab < nab: the FP64-accumulator seed of each anchor block for the
// spin of row slot gi % 16.  The basis is evaluated once per spin.
__device__ inline void seed_elem(int64_t gi, int64_t ns, int64_t ngroups, int64_t nab,
                                 const PhaseSplit &ps, const double *x, const double *y,
                                 const double *z, const double *dw, const double *th0,
                                 const double *anc, double *seed) {
    const int64_t s = (gi / 16) * 16 + slot_spin((int)(gi % 16));
    double b[12];
    for (int q = 0; q < 12; ++q)
        b[q] = (s < ns && q < ps.nho) ? basis_term(ps.codeho[q], x[s], y[s], z[s], dw[s]) : 0.0;
    const double t0 = th0[gi];
    // Fixed trip count, unconditional loads from a clamped index (b[q] is 0
    // for q >= nho, and fma(0, a, v) = v): b stays in registers and the
    // anchors of one block are loaded together.  A runtime-bound loop indexed
    // b indirectly and waited for each scalar load in turn.
    const int qmax = ps.nho > 0 ? ps.nho - 1 : 0;
    for (int64_t ab = 0; ab < nab; ++ab) {
        double v = t0;
        if (ps.nho > 0) {
#pragma unroll
            for (int q = 0; q < 12; ++q) v = fma(b[q], anc[ab * ps.nho + (q < qmax ? q : qmax)], v);
        }
        seed[ab * ngroups * 16 + gi] = v;
    }
}

// idx < nadc * nc, coil fastest so a wave reads contiguous partial rows; FP64
// sum over spin tiles in tile order, scales, then the spin-independent
// zeroth-order phase in FP64.
__device__ inline void finalize_elem(int64_t idx, int64_t nc, int64_t nstiles, int64_t rows,
                                     int64_t cols, const float *partials, float m0max,
                                     const double *csc, const double *kr0, double2 *out,
                                     int64_t ld) {
    const int64_t c = idx % nc, k = idx / nc;
    double re = 0.0, im = 0.0;
    // unrolled so that eight loads are in flight; the sum stays in tile order
#pragma unroll 8
    for (int64_t st = 0; st < nstiles; ++st) {
        const float2 p = *reinterpret_cast<const float2 *>(partials + (st * rows + k) * cols + 2 * c);
        re += p.x;
        im += p.y;
    }
    const double sc = ldexp(col_scale(m0max) * csc[c], -2 * kOpScale);
    re *= sc;
    im *= sc;
    const double r = kr0[k] - floor(kr0[k]);
    double sn, cs;
    sincos(kTwoPi * r, &sn, &cs);
    out[k + c * ld] = double2{re * cs - im * sn, re * sn + im * cs};
}

// s < ns: FP64 end-of-segment state, Mxy <- M0 e^{-T/T2} e^{i phi_end},
// Mz <- Mz e^{-T/T1} + rho (1 - e^{-T/T1}).
__device__ inline void end_elem(int64_t s, const TermList &tl, const double *x, const double *y,
                                const double *z, const double *dw, const double *rho,
                                const double *T1, const double *T2, const double *kend, double T,
                                double2 *mxy, double *mz) {
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

// ---- fused phase + state + coil GEMM ----

// State element (d cos, d sin) of one spin-sample, split into FP16 halves:
// hi = fp16(x) as one opaque packed register, lo = fp16(x - hi) from it.
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

// Operands of one main-kernel pass: one segment chunk (tiles [tile_lo,
// tile_hi) of the segment tables, sample blocks from sb_lo) and one range of
// coil blocks [cb0, cb0 + NCB) of ncb_total.  seed holds anchor blocks from
// ab_lo (with 16 values of padding after the last); basis64 has ngroups + 1
// groups; partials hold the chunk's rows only.
struct MainOps {
    int64_t ngroups;
    int gpt;
    int cb0, ncb_total;
    int64_t tile_lo, tile_hi, sb_lo, ab_lo;
    const double *basis64;
    const half8_t *aho;
    const float *cdec;
    const float *lm;
    const double *seed;
    const half8_t *bop;
    const double *kref;   // per tile: kRefsPerTile references x (x, y, z, dw) Kr, FP64
    const half8_t *bho;
    const float *tk;
    float *partials;
    const double *basisA;  // all-FP64 body
    const double *th0;
    const double *kr;
};

// Work item of one workgroup: spin tile and chunk-relative sample block.
struct WorkItem {
    int64_t stile, tile0, anchor;
    bool valid;
};

__device__ inline WorkItem make_item(const MainOps &o, int64_t sblock, int64_t stile) {
    const int w = threadIdx.x >> 6;
    const int64_t sb = o.sb_lo + sblock;
    WorkItem it;
    it.stile = stile;
    it.tile0 = (sb * kWaves + w) * kTiles;
    it.anchor = sb / kBlocksPerAnchor - o.ab_lo;
    it.valid = it.tile0 < o.tile_hi;
    return it;
}

template <int NCB>
__device__ inline void write_partials(const floatx4 (&acc)[kTiles][NCB], const WorkItem &it,
                                      const MainOps &o) {
    const int lane = threadIdx.x & 63;
    const int64_t rows = (o.tile_hi - o.tile_lo) * 16, cols = 16 * (int64_t)o.ncb_total;
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        const int64_t tile = it.tile0 + t;
        if (tile >= o.tile_hi) continue;
#pragma unroll
        for (int c = 0; c < NCB; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t row = (tile - o.tile_lo) * 16 + 4 * (lane >> 4) + v;
                const int64_t col = 16 * (o.cb0 + c) + (lane & 15);
                o.partials[(it.stile * rows + row) * cols + col] = acc[t][c][v];
            }
    }
}

// Source of one 16-byte chunk of a group record: chunk c of group grp is at
// p + grp * stride.  Record layout: coil B operand (nbop chunks), offset A
// operand (naho), decay rate (4), log-magnitude (4), then the next group's
// FP64 first-order basis (32) and phase seed (8), from which one wave
// computes the next group's first-order anchors.  Chunks at or beyond nrec
// are never loaded.
struct StageSrc {
    const unsigned char *p;
    int64_t stride;
};

template <int NMH>
__device__ inline StageSrc stage_src(int c, int nbop, int naho, int nrec, const MainOps &o,
                                     const double *seed_b) {
    if (c >= nrec) c = 0;
    if (c < nbop)
        return {reinterpret_cast<const unsigned char *>(o.bop) + ((int64_t)o.cb0 * 2 * 64 + c) * 16,
                (int64_t)o.ncb_total * 2 * 64 * 16};
    if ((c -= nbop) < naho)
        return {reinterpret_cast<const unsigned char *>(o.aho) + (int64_t)c * 16, (int64_t)NMH * 64 * 16};
    if ((c -= naho) < 4) return {reinterpret_cast<const unsigned char *>(o.cdec) + (int64_t)c * 16, 16 * 4};
    if ((c -= 4) < 4) return {reinterpret_cast<const unsigned char *>(o.lm) + (int64_t)c * 16, 16 * 4};
    if ((c -= 4) < 32)  // group grp + 1
        return {reinterpret_cast<const unsigned char *>(o.basis64) + 64 * 8 + (int64_t)c * 16, 64 * 8};
    c -= 32;
    return {reinterpret_cast<const unsigned char *>(seed_b) + 16 * 8 + (int64_t)c * 16, 16 * 8};
}

// LDS of the split-phase body: two group-record buffers (double buffer) and
// two first-order anchor buffers (16 references x 16 spins, FP32).  They are
// separate __shared__ arrays so the compiler can tell that the LDS accesses
// of one buffer do not alias the direct global-to-LDS copy into the other
// record, and does not wait for that copy before reading.
template <int NMH, int NCB>
struct SplitLds {
    static constexpr int kBopChunks = NCB * 2 * 64;    // 16-byte chunks
    static constexpr int kAhoChunks = NMH * 64;
    static constexpr int kTailChunks = 4 + 4;          // decay rate, log-magnitude
    static constexpr int kRecChunks = kBopChunks + kAhoChunks + kTailChunks + 32 + 8;
    static constexpr int kRecBytes = kRecChunks * 16;
    static constexpr int kAncBytes = 16 * 16 * 4;
    static constexpr int kStage = (kRecChunks + 255) / 256;  // chunks per thread
};

struct SplitSmem {
    unsigned char *rec0, *rec1;  // group records, SplitLds::kRecBytes each
    float *anc0, *anc1;          // first-order anchors of the records' groups
};

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void global_cvoid_t;

// Direct 16-byte global-to-LDS copy (global_load_lds_dwordx4): lane l of the
// wave writes lds_wave + 16 l; lds_wave must be the same for the whole wave.
__device__ inline void copy16_to_lds(const unsigned char *src, unsigned char *lds_wave) {
    __builtin_amdgcn_global_load_lds((global_cvoid_t *)src, (lds_void_t *)lds_wave, 16, 0, 0);
}

// Wait until this wave's global loads, and with them its LDS copies, are done.
__device__ inline void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)

// Split-phase body (FP64 anchors + split-FP16 offsets) for one work item.
// Workgroup = kWaves waves over one spin tile; wave w owns kTiles consecutive
// 16-sample tiles and accumulates a 16-sample x (16 NCB)-column FP32 tile per
// sample tile over every 16-spin group of the spin tile.
//
// The spin-group operands (coil B operand, offset A operand, decay rate,
// log-magnitude) are the same for all waves: the workgroup copies each
// group's record into LDS once, double-buffered, with direct global-to-LDS
// loads issued before the compute of the previous group and awaited after it,
// one barrier per group.  Groups go in pairs, the even one from rec0/anc0 and
// the odd one from rec1/anc1.  A record also carries the next group's FP64
// first-order basis and phase seed: while the workgroup computes group g, one
// wave (rotating) turns them into group g + 1's first-order anchors for the
// 16 tile references of the sample block (one FP64 MFMA) and writes them
// to the other anchor buffer, which the barrier publishes.  The per-tile
// offset B operands stay in registers for the whole work item (reading
// them from LDS every group would nearly double the group loop's LDS
// reads; fewer LDS bytes per spin-sample measured faster).  Every wave
// takes part in the copies and barriers, with or without tiles; the body ends
// on a barrier, so a persistent caller may reuse the LDS at once.
template <int NMH, int NCB>
__device__ __forceinline__ void main_split_body(const MainOps &o, const WorkItem &it,
                                                const SplitSmem &sm) {
    static_assert(NMH >= 1, "the first-order offsets need at least one FP16 MFMA");
    using L = SplitLds<NMH, NCB>;
    // Opaque thread id: a persistent caller runs this body in a loop, and
    // lane-dependent addresses must not be hoisted out of it (they would stay
    // live across items and push the kernel past its register budget).
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63;
    const int w = tid >> 6;
    const bool active = it.valid;

    float tkv[kTiles];
    half8_t bvr[kTiles][NMH];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        const int64_t tile = it.tile0 + t;
        const bool ok = active && tile < o.tile_hi;
        tkv[t] = ok ? o.tk[tile * 16 + (lane & 15)] : 0.0f;
#pragma unroll
        for (int m = 0; m < NMH; ++m)
            bvr[t][m] = ok ? o.bho[(tile * NMH + m) * 64 + lane] : half8_t{};
    }
    // B operand of the anchor MFMA: lane l holds Kr of first-order term l/16
    // at reference column l%16 (tile wtile0 + col/2 of the sample block, half
    // col%2).
    const int64_t wtile0 = it.tile0 - (int64_t)w * kTiles;
    double kref;
    {
        const int col = lane & 15;
        const int64_t tile = wtile0 + col / kRefsPerTile;
        kref = tile < o.tile_hi ? o.kref[(wtile0 * kRefsPerTile + col) * 4 + (lane >> 4)] : 0.0;
    }

    // Each thread copies chunks tid, tid + 256 and tid + 512 of every group
    // record; wave w's lanes write the contiguous chunks 64 w + 256 k + lane.
    static_assert(L::kStage <= 3, "a group record has at most 768 chunks");
    const double *seed_b = o.seed + it.anchor * o.ngroups * 16;
    const StageSrc s0 = stage_src<NMH>(tid, L::kBopChunks, L::kAhoChunks, L::kRecChunks, o, seed_b);
    const StageSrc s1 = stage_src<NMH>(tid + 256, L::kBopChunks, L::kAhoChunks, L::kRecChunks, o, seed_b);
    const StageSrc s2 = stage_src<NMH>(tid + 512, L::kBopChunks, L::kAhoChunks, L::kRecChunks, o, seed_b);
    const bool u0 = tid < L::kRecChunks, u1 = tid + 256 < L::kRecChunks, u2 = tid + 512 < L::kRecChunks;
    // The wave's LDS offset as a scalar (the copy's M0 base is then formed
    // with scalar adds, not a readfirstlane per copy), and no lane mask for
    // the copies that every thread makes (1.3 % faster at 32 coils).
    const int wbase = __builtin_amdgcn_readfirstlane((tid & ~63) * 16);
    // The loop below stages the groups g0, g0 + 1, g0 + 2, ... in that order,
    // each once, so the sources are running pointers advanced by one record
    // stride per call: a 64-bit add instead of a per-lane 64-bit multiply-add
    // per copy, and 213 instead of 225 VGPRs.  Two waves per SIMD then leave
    // 80 VGPRs, room for the 48-64-VGPR kernels of the plan's other stream
    // to run beside this one instead of waiting for its workgroups to end
    // (about 1 % per step at 1M spins x 32 coils).
    const int64_t g0 = it.stile * o.gpt;
    const int64_t g1 = std::min<int64_t>(g0 + o.gpt, o.ngroups);
    const unsigned char *q0 = s0.p + g0 * s0.stride, *q1 = s1.p + g0 * s1.stride, *q2 = s2.p + g0 * s2.stride;
    auto stage = [&](unsigned char *dst) {
        if (L::kRecChunks >= 256 || u0) copy16_to_lds(q0, dst + wbase);
        if (L::kStage > 1 && (L::kRecChunks >= 512 || u1)) copy16_to_lds(q1, dst + wbase + 256 * 16);
        if (L::kStage > 2 && u2) copy16_to_lds(q2, dst + wbase + 512 * 16);
        q0 += s0.stride;
        q1 += s1.stride;
        q2 += s2.stride;
    };

    // First-order anchors of one group from its FP64 basis (A: spin l%16,
    // term l/16) and seed (C: spins l/16 + 4v, as the output): lane l writes
    // the four spins' anchors at reference column l%16, reduced to [0, 1).
    auto anchors = [&](double ba, doublex4 sd, float *dst) {
        const doublex4 d = __builtin_amdgcn_mfma_f64_16x16x4f64(ba, kref, sd, 0, 0, 0);
        floatx4 a;
#pragma unroll
        for (int v = 0; v < 4; ++v) a[v] = reduce_rev(d[v]);
        reinterpret_cast<floatx4 *>(dst)[(lane & 15) * 4 + (lane >> 4)] = a;
    };

    floatx4 acc[kTiles][NCB];
#pragma unroll
    for (int t = 0; t < kTiles; ++t)
#pragma unroll
        for (int c = 0; c < NCB; ++c) acc[t][c] = floatx4{0.f, 0.f, 0.f, 0.f};

    // Anchor column read by this lane for local tile t: the half of lane%16.
    const int acol0 = (w * kTiles) * kRefsPerTile + (lane & 15) / kRefHalf;

    auto compute = [&](const unsigned char *rec, const float *anc, float *anc_next, bool next) {
        const half8_t *rb = reinterpret_cast<const half8_t *>(rec);
        const unsigned char *tail = rec + (L::kBopChunks + L::kAhoChunks) * 16;
        if (next) {
            // Two 16-byte reads: a 32-byte LDS read would make the compiler
            // wait for the in-flight copy into the other record buffer.
            const unsigned char *nx = tail + L::kTailChunks * 16;
            const double ba = reinterpret_cast<const double *>(nx)[lane];
            const double2 sd01 = reinterpret_cast<const double2 *>(nx + 512)[2 * (lane >> 4)];
            const double2 sd23 = reinterpret_cast<const double2 *>(nx + 512)[2 * (lane >> 4) + 1];
            anchors(ba, doublex4{sd01.x, sd01.y, sd23.x, sd23.y}, anc_next);
        }
        if (!active) return;  // a wave past the segment's last tile
        half8_t bh[NCB], bl[NCB];
#pragma unroll
        for (int c = 0; c < NCB; ++c) {
            bh[c] = rb[(c * 2 + 0) * 64 + lane];
            bl[c] = rb[(c * 2 + 1) * 64 + lane];
        }
        half8_t ah[NMH];
#pragma unroll
        for (int m = 0; m < NMH; ++m) ah[m] = rb[L::kBopChunks + m * 64 + lane];
        const floatx4 cd = reinterpret_cast<const floatx4 *>(tail)[lane >> 4];
        const floatx4 lmv = reinterpret_cast<const floatx4 *>(tail + 64)[lane >> 4];
        // Offset MFMAs of tile t (phase offsets of its 16 x 16 spin-samples).
        // Tiles 0 and 1 go first; tile t + 2's are issued among tile t's coil
        // MFMAs, so every tile's offsets are ready two tiles before its
        // states need them (1.3-1.6 % faster than issuing all four first).
        floatx4 offt[kTiles];
        auto offsets = [&](int t) {
            offt[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int m = 0; m < NMH; ++m)
                offt[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[m], bvr[t][m], offt[t], 0, 0, 0);
        };
        offsets(0);
        if (kTiles > 1) offsets(1);
        // FP16 halves of the states of tile t: phase = anchor + offset, then
        // sin, cos and decay
        auto states = [&](int t, half8_t &ahi, half8_t &alo) {
            const floatx4 an =
                reinterpret_cast<const floatx4 *>(anc)[(acol0 + t * kRefsPerTile) * 4 + (lane >> 4)];
            const floatx4 off = offt[t];
            uint4 ahu, alu;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float rf = fmaf(off[v], 0x1p-20f, an[v]);
                const StateSplit e = state_split(rf, cd[v], tkv[t], lmv[v]);
                (&ahu.x)[v] = __builtin_bit_cast(unsigned int, e.hi);
                (&alu.x)[v] = __builtin_bit_cast(unsigned int, e.lo);
            }
            ahi = __builtin_bit_cast(half8_t, ahu);
            alo = __builtin_bit_cast(half8_t, alu);
        };
        // Software pipeline: the states of tile t + 1 are computed while the
        // coil MFMAs of tile t (and the offset MFMAs of tile t + 2) issue,
        // one MFMA per four VALU instructions, per three when the offset
        // MFMAs join (measured 2 % faster than tile by tile; a 1:2
        // interleave, and the pipeline without one, gained less).
        half8_t sh[2], sl[2];
        states(0, sh[0], sl[0]);
#pragma unroll
        for (int t = 0; t < kTiles; ++t) {
            if (t + 2 < kTiles) offsets(t + 2);
            if (t + 1 < kTiles) states(t + 1, sh[(t + 1) & 1], sl[(t + 1) & 1]);
            const half8_t ahi = sh[t & 1], alo = sl[t & 1];
#pragma unroll
            for (int c = 0; c < NCB; ++c) {
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bh[c], acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bl[c], acc[t][c], 0, 0, 0);
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, bh[c], acc[t][c], 0, 0, 0);
            }
            if (t + 2 < kTiles) {
#pragma unroll
                for (int k = 0; k < 3 * NCB + NMH; ++k) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
                    __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);  // three VALU
                }
            } else if (t + 1 < kTiles) {
#pragma unroll
                for (int k = 0; k < 3 * NCB; ++k) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
                    __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);  // four VALU
                }
            }
        }
    };

    stage(sm.rec0);  // group g0
    if (w == 0) {
        // anchors of the first group (later groups' come with the records)
        const double ba = o.basis64[g0 * 64 + lane];
        const doublex4 sd = *reinterpret_cast<const doublex4 *>(seed_b + g0 * 16 + 4 * (lane >> 4));
        anchors(ba, sd, sm.anc0);
    }
    wait_vmem();
    __syncthreads();
    // wave computing the anchors of group grp + 1 during group grp
    auto next_wave = [&](int64_t grp) { return grp + 1 < g1 && w == (int)((grp + 1 - g0) & (kWaves - 1)); };
    for (int64_t grp = g0; grp < g1; grp += 2) {
        if (grp + 1 < g1) stage(sm.rec1);  // group grp + 1
        compute(sm.rec0, sm.anc0, sm.anc1, next_wave(grp));
        wait_vmem();
        __syncthreads();
        if (grp + 1 >= g1) break;
        if (grp + 2 < g1) stage(sm.rec0);  // group grp + 2
        compute(sm.rec1, sm.anc1, sm.anc0, next_wave(grp + 1));
        wait_vmem();
        __syncthreads();
    }
    if (active) write_partials<NCB>(acc, it, o);
}

// All-FP64 phase body (NI FP64 MFMAs over every term, accumulator seeded
// with arg(M0)), used when a segment's higher-order offsets do not fit FP16.
// No LDS and no barriers.
template <int NI, int NCB>
__device__ __forceinline__ void main64_body(const MainOps &o, const WorkItem &it) {
    if (!it.valid) return;
    const int lane = threadIdx.x & 63;
    double krv[kTiles][NI];
    float tkv[kTiles];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) {
        const int64_t tile = it.tile0 + t;
        const bool ok = tile < o.tile_hi;
#pragma unroll
        for (int i = 0; i < NI; ++i) krv[t][i] = ok ? o.kr[(tile * NI + i) * 64 + lane] : 0.0;
        tkv[t] = ok ? o.tk[tile * 16 + (lane & 15)] : 0.0f;
    }
    floatx4 acc[kTiles][NCB];
#pragma unroll
    for (int t = 0; t < kTiles; ++t)
#pragma unroll
        for (int c = 0; c < NCB; ++c) acc[t][c] = floatx4{0.f, 0.f, 0.f, 0.f};

    const int64_t g0 = it.stile * o.gpt;
    const int64_t g1 = std::min<int64_t>(g0 + o.gpt, o.ngroups);
    for (int64_t grp = g0; grp < g1; ++grp) {
        double ba[NI];
#pragma unroll
        for (int i = 0; i < NI; ++i) ba[i] = o.basisA[(grp * NI + i) * 64 + lane];
        const floatx4 cd = *reinterpret_cast<const floatx4 *>(o.cdec + grp * 16 + 4 * (lane >> 4));
        const floatx4 lmv = *reinterpret_cast<const floatx4 *>(o.lm + grp * 16 + 4 * (lane >> 4));
        const doublex4 ph0 = *reinterpret_cast<const doublex4 *>(o.th0 + grp * 16 + 4 * (lane >> 4));
        half8_t bh[NCB], bl[NCB];
#pragma unroll
        for (int c = 0; c < NCB; ++c) {
            bh[c] = o.bop[((grp * o.ncb_total + o.cb0 + c) * 2 + 0) * 64 + lane];
            bl[c] = o.bop[((grp * o.ncb_total + o.cb0 + c) * 2 + 1) * 64 + lane];
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
    write_partials<NCB>(acc, it, o);
}

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace hom
