// mixed_kernels.hip -- per-plan launches of the segment-fused mixed-precision
// BlochHighOrder path (math, layouts and device bodies in mixed_common.h).
//
// Cached per (phantom, CSM, mask):
//   k_mix_basis64 / k_mix_basis  FP64 basis values in FP64-MFMA A-operand order
//   k_mix_aho      split-FP16 basis of the phase offsets (two-way higher
//                  order, three-way first order), FP16-MFMA A-operand order
//   k_mix_cdec     -log2(e)/T2 per spin, in MFMA output-row order
//   k_mix_bop      B operand of the coil GEMM: split-FP16 (hi, lo) of
//                  csm / S_c (S_c a power-of-two column scale)
// Per segment:
//   k_mix_m0max    max |M0| over spins (power-of-two state scale S)
//   k_mix_seg      per spin: arg(M0) in revolutions (FP64), log2(|M0|/S) + 14
//   k_mix_seed     per (anchor block, spin): arg(M0)/2pi + higher-order anchor
//   k_mix_main     per 16-spin group and 256-sample block: first-order FP64
//                  anchors at 16 tile references (one FP64 MFMA); per 16
//                  spins x 16 samples: offsets (split-FP16 MFMA), anchor +
//                  offset, v_sin/v_cos
//                  (revolutions), v_exp for |M0| e^{-t/T2}, FP16 split of the
//                  state, complex coil GEMM as three split-FP16 products on
//                  v_mfma_f32_16x16x32_f16, FP32 accumulation over a spin tile
//   k_mix_finalize FP64 sum of the spin-tile partials in tile order, scales
//                  and the spin-independent zeroth-order phase in FP64
//   k_mix_end      FP64 end-of-segment state: Mxy <- M0 e^{-T/T2} e^{i phi_end},
//                  Mz <- Mz e^{-T/T1} + rho (1 - e^{-T/T1})
// RF blocks run the f64 per-block excitation kernel unchanged.
#include "mixed_common.h"

#include <cstring>
#include <system_error>
#include <thread>

namespace hom {

struct MixedWorkspace {
    int basis_flags = -1;
    int table_flags = -1;
    std::vector<int64_t> table_parts;

    TermList terms{};
    PhaseSplit split{};
    int NI = 0, NCB = 0;
    int64_t ngroups = 0;
    double *basisA = nullptr;   // ngroups x NI x 64 (all-FP64 fallback)
    double *basis64 = nullptr;  // (ngroups + 1) x 64
    half8_t *aho = nullptr;     // ngroups x nmh x 64
    float *cdec = nullptr;      // ngroups x 16, MFMA output-row order
    half8_t *bop = nullptr;     // ngroups x NCB x 2 x 64
    double *csc = nullptr;      // 8 * NCB power-of-two column scales
    // Per-segment operands in two slots (segment k with samples uses slot
    // k % 2, so that the small kernels of one segment run beside the main
    // kernel of the previous one):
    double *th0 = nullptr;      // 2 x ngroups x 16: arg(M0) / 2pi
    float *lm = nullptr;        // 2 x ngroups x 16: log2(|M0| / S) + kOpScale
    float *m0max = nullptr;     // 2: max |M0| of the segment
    double *seed = nullptr;     // 2 x (max_nab x ngroups x 16 + 16 padding)
    int64_t seed_cap = 0;       // per slot
    double *kr = nullptr;       // fallback FP64 tables (all terms)
    double *kref = nullptr;     // FP64 Kr of (x, y, z, dw) at the tile references
    half8_t *bho = nullptr;     // FP16 offset tables
    double *anc = nullptr;      // per segment, per anchor block: higher-order anchors
    float *tk = nullptr;        // per-sample time within segment
    double *kr0 = nullptr;      // per-sample zeroth-order phase (revolutions)
    double *kend = nullptr;     // per segment: all terms, then kr0_end
    float *partials = nullptr;  // 2 slots
    int64_t partial_cap = 0;    // per slot
    std::vector<Segment> segs;
    // The main kernels' stream, and the events that order it against the
    // plan's stream (prep done / main done per slot; the end-of-run join).
    hipStream_t main_stream = nullptr;
    hipEvent_t prep_done[2] = {nullptr, nullptr}, main_done[2] = {nullptr, nullptr}, joined = nullptr;
};

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
    s.n1 = 0;
    for (int q = 0; q < 4; ++q) {
        s.code1[q] = 0;
        s.aexp1[q] = 0;
    }
    for (int q = 0; q < 4; ++q) {
        if (!s.code64[q]) continue;
        const int c = s.code64[q];
        s.code1[s.n1] = c;
        if (basis_max[c] > 0.0) {
            int e = 0;
            frexp(basis_max[c], &e);
            s.aexp1[s.n1] = kFOA - e;  // max |A| in [2^13, 2^14)
        }
        ++s.n1;
    }
    s.nmh = (3 * s.nho + kFOParts * s.n1 + 31) / 32;
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

namespace {

__global__ void k_mix_basis(int64_t ns, int64_t ngroups, int NI, TermList tl,
                            const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, const double *__restrict__ dw,
                            double *__restrict__ basisA) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < ngroups * NI * 64) basisA_elem(idx, ns, NI, tl, x, y, z, dw, basisA);
}

__global__ void k_mix_basis64(int64_t ns, int64_t ngroups, PhaseSplit ps,
                              const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ z, const double *__restrict__ dw,
                              double *__restrict__ basis64) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < (ngroups + 1) * 64) basis64_elem(idx, ns, ps, x, y, z, dw, basis64);
}

__global__ void k_mix_aho(int64_t ns, int64_t ngroups, PhaseSplit ps,
                          const double *__restrict__ x, const double *__restrict__ y,
                          const double *__restrict__ z, const double *__restrict__ dw,
                          half8_t *__restrict__ aho) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < ngroups * ps.nmh * 64) aho_elem(idx, ns, ps, x, y, z, dw, aho);
}

__global__ void k_mix_cdec(int64_t ns, int64_t ngroups, const double *__restrict__ T2,
                           float *__restrict__ cdec) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < ngroups * 16) cdec_elem(idx, ns, T2, cdec);
}

__global__ void k_fill_f32(float *a, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_mix_bop(int64_t ns, int64_t nc, int64_t ngroups, int NCB,
                          const double2 *__restrict__ csm, int64_t ldc,
                          const double *__restrict__ csc, half8_t *__restrict__ bop) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < ngroups * NCB * 64) bop_elem(idx, ns, nc, NCB, csm, ldc, csc, bop);
}

// Grid-stride over the spins with a capped grid: one atomic per workgroup
// (one per 256 spins serialised 4096 atomics on one word at 1M spins).
__global__ __launch_bounds__(256) void k_mix_m0max(int64_t ns, const double2 *__restrict__ mxy,
                                                   float *__restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    float v = 0.0f;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += stride) {
        const double2 m = mxy[s];
        v = fmaxf(v, (float)sqrt(m.x * m.x + m.y * m.y));
    }
    m0max_reduce(v, out);
}

__global__ void k_mix_seg(int64_t ns, int64_t ngroups, const double2 *__restrict__ mxy,
                          const float *__restrict__ m0max, double *__restrict__ th0,
                          float *__restrict__ lm) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < ngroups * 16) seg_elem(idx, ns, mxy, *m0max, th0, lm);
}

__global__ void k_mix_seed(int64_t ns, int64_t ngroups, int64_t nab, PhaseSplit ps,
                           const double *__restrict__ x, const double *__restrict__ y,
                           const double *__restrict__ z, const double *__restrict__ dw,
                           const double *__restrict__ th0, const double *__restrict__ anc,
                           double *__restrict__ seed) {
    const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi < ngroups * 16) seed_elem(gi, ns, ngroups, nab, ps, x, y, z, dw, th0, anc, seed);
}

// Hardware places block b on XCD b % 8; the remap gives each XCD a contiguous
// run of work items, and consecutive items share a spin tile, so its operands
// are reused from that XCD's L2.
__device__ inline int64_t xcd_work() {
    const int64_t nb = gridDim.x, bid = blockIdx.x;
    const int64_t per = nb / 8, rem = nb % 8, xcd = bid % 8;
    return xcd * per + (xcd < rem ? xcd : rem) + bid / 8;
}

// One work item (spin tile x sample block) per workgroup; nsb sample blocks.
template <int NMH, int NCB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HOM_WAVES_PER_EU, HOM_WAVES_PER_EU))) void k_mix_main(
    MainOps o, int64_t nsb, int64_t n_work) {
    using L = SplitLds<NMH, NCB>;
    __shared__ __attribute__((aligned(32))) unsigned char rec0[L::kRecBytes];
    __shared__ __attribute__((aligned(32))) unsigned char rec1[L::kRecBytes];
    __shared__ __attribute__((aligned(32))) float anc0[L::kAncBytes / 4];
    __shared__ __attribute__((aligned(32))) float anc1[L::kAncBytes / 4];
    const int64_t work = xcd_work();
    if (work >= n_work) return;  // whole workgroup
    main_split_body<NMH, NCB>(o, make_item(o, work % nsb, work / nsb), SplitSmem{rec0, rec1, anc0, anc1});
}

template <int NI, int NCB>
__global__ __launch_bounds__(256) void k_mix_main64(MainOps o, int64_t nsb, int64_t n_work) {
    const int64_t work = xcd_work();
    if (work >= n_work) return;
    main64_body<NI, NCB>(o, make_item(o, work % nsb, work / nsb));
}

__global__ void k_mix_finalize(int64_t nadc, int64_t nc, int64_t nstiles, int64_t rows, int64_t cols,
                               const float *__restrict__ partials, const float *__restrict__ m0max,
                               const double *__restrict__ csc, const double *__restrict__ kr0,
                               double2 *__restrict__ out, int64_t ld) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nadc * nc)
        finalize_elem(idx, nc, nstiles, rows, cols, partials, *m0max, csc, kr0, out, ld);
}

__global__ void k_mix_end(int64_t ns, TermList tl, const double *__restrict__ x,
                          const double *__restrict__ y, const double *__restrict__ z,
                          const double *__restrict__ dw, const double *__restrict__ rho,
                          const double *__restrict__ T1, const double *__restrict__ T2,
                          const double *__restrict__ kend, double T, double2 *__restrict__ mxy,
                          double *__restrict__ mz) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns) end_elem(s, tl, x, y, z, dw, rho, T1, T2, kend, T, mxy, mz);
}

template <int NMH>
hipError_t launch_split_nmh(int ncb, dim3 grid, hipStream_t st, const MainOps &o, int64_t nsb,
                            int64_t nwork) {
#define HOM_MAIN(NCBV)                                                                       \
    hipLaunchKernelGGL((k_mix_main<NMH, NCBV>), grid, dim3(256), 0, st, o, nsb, nwork); \
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
hipError_t launch_fp64_ni(int ncb, dim3 grid, hipStream_t st, const MainOps &o, int64_t nsb,
                          int64_t nwork) {
#define HOM_MAIN(NCBV)                                                                         \
    hipLaunchKernelGGL((k_mix_main64<NI, NCBV>), grid, dim3(256), 0, st, o, nsb, nwork); \
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

hipError_t launch_main(int ncb, bool fallback, int nmh, int NI, hipStream_t st, const MainOps &o,
                       int64_t nsb, int64_t nwork) {
    const dim3 grid((unsigned)nwork);
    if (!fallback) {
        switch (nmh) {
            case 1: return launch_split_nmh<1>(ncb, grid, st, o, nsb, nwork);
            case 2: return launch_split_nmh<2>(ncb, grid, st, o, nsb, nwork);
            default: return hipErrorInvalidValue;
        }
    }
    switch (NI) {
        case 1: return launch_fp64_ni<1>(ncb, grid, st, o, nsb, nwork);
        case 2: return launch_fp64_ni<2>(ncb, grid, st, o, nsb, nwork);
        case 3: return launch_fp64_ni<3>(ncb, grid, st, o, nsb, nwork);
        case 4: return launch_fp64_ni<4>(ncb, grid, st, o, nsb, nwork);
        default: return hipErrorInvalidValue;
    }
}

// Replaces *dst with a device buffer of n elements (none for n = 0).
template <typename T>
int dev_alloc(hom_plan *p, T **dst, size_t n) {
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    if (n == 0) return HOM_OK;
    if (hom::dev_malloc((void **)dst, sizeof(T) * n) != hipSuccess) {
        *dst = nullptr;
        return set_error(p, HOM_ERR_MEMORY, "device allocation failed for mixed-path tables");
    }
    return HOM_OK;
}

// Copies elements [off, off + n) of a host table to the same range of its
// device buffer (synchronous; the plan's stream may still run earlier
// segments, which read other ranges).
template <typename T>
int upload_range(hom_plan *p, T *dev, const std::vector<T> &host, int64_t off, int64_t n) {
    if (n <= 0) return HOM_OK;
    return check_hip(p, hipMemcpy(dev + off, host.data() + off, sizeof(T) * (size_t)n, hipMemcpyHostToDevice),
                     "hipMemcpy");
}

// Nearest-even double -> binary16 bits (inf on overflow, signed zero on
// underflow), inline integer arithmetic in place of the runtime's conversion
// call in the per-sample table loop.
inline uint16_t half_bits_rne(double v) {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    const uint16_t sign = (uint16_t)((b >> 48) & 0x8000u);
    const int e = (int)((b >> 52) & 0x7ff);
    const uint64_t frac = b & 0xFFFFFFFFFFFFFull;
    if (e == 0x7ff) return (uint16_t)(sign | 0x7c00u | (frac ? 0x200u : 0u));
    if (e == 0) return sign;  // double subnormals lie far below half's range
    const int E = e - 1023;
    if (E > 15) return (uint16_t)(sign | 0x7c00u);
    int Eh = E < -14 ? -14 : E;
    const int shift = 42 + (Eh - E);  // drops the bits below the half ulp 2^(Eh-10)
    if (shift >= 64) return sign;
    const uint64_t sig = frac | (1ull << 52);
    uint64_t q = sig >> shift;
    const uint64_t rem = sig & ((1ull << shift) - 1), halfway = 1ull << (shift - 1);
    if (rem > halfway || (rem == halfway && (q & 1))) ++q;
    if (q >= 2048) {
        q >>= 1;
        ++Eh;
    }
    if (q < 1024) return (uint16_t)(sign | q);  // subnormal
    if (Eh > 15) return (uint16_t)(sign | 0x7c00u);
    return (uint16_t)(sign | ((Eh + 15) << 10) | (q - 1024));
}

// Exact binary16 bits -> double.
inline double half_bits_to_double(uint16_t h) {
    const int ex = (h >> 10) & 31, m = h & 1023;
    double v;
    if (ex == 0) {
        v = (double)m * 0x1p-24;
    } else if (ex == 31) {
        v = m ? std::nan("") : HUGE_VAL;
    } else {
        const uint64_t b = (uint64_t)(ex - 15 + 1023) << 52 | (uint64_t)m << 42;
        std::memcpy(&v, &b, 8);
    }
    return (h & 0x8000) ? -v : v;
}

// v ~= hi + lo with hi = fp16(v) and lo = fp16(v - hi), each a correctly
// rounded conversion from double.
void host_split16(double v, _Float16 &hi, _Float16 &lo) {
    const uint16_t hb = half_bits_rne(v);
    const uint16_t lb = half_bits_rne(v - half_bits_to_double(hb));
    std::memcpy(&hi, &hb, 2);
    std::memcpy(&lo, &lb, 2);
}

// v ~= p[0] + p[1] + p[2], each part the correctly rounded fp16 of what the
// previous ones leave (the differences are exact in FP64).
void host_split16x3(double v, _Float16 (&p)[3]) {
    for (int k = 0; k < 3; ++k) {
        const uint16_t b = half_bits_rne(v);
        std::memcpy(&p[k], &b, 2);
        v -= half_bits_to_double(b);
    }
}

template <typename T>
void free_dev(T *&q) {
    if (q) (void)hipFree(q);
    q = nullptr;
}

}  // namespace

// Host FP64 integration of the coefficient waveforms per segment, written in
// MFMA B-operand lane order:
//   kr   (all-FP64 fallback): tile tau, step i, lane l -> term 4i + l/16 at sample l%16
//   kref (first-order anchors): tile tau, half h, term q of (x, y, z, dw) ->
//        Kr_q at the half's reference sample (ref_sample)
//   bho  (FP16 offsets):      tile tau, MFMA m, lane l, element i -> K slot
//                             32m + 8(l/16) + i at sample l%16: higher orders
//                             from the 256-sample anchor, first order from the
//                             sample's tile reference
//   anc: per anchor block, the higher-order Kr at the block's first sample.
// Runs fill(k0, k1, max) over [0, n) in up to 8 host threads (chunks of at
// least 2048 items) and returns the largest of the threads' maxima.
template <typename F>
double parallel_max(int64_t n, F &&fill) {
    const int64_t hw = std::max<int64_t>(1, (int64_t)std::thread::hardware_concurrency());
    const int64_t nth = std::max<int64_t>(1, std::min<int64_t>({8, hw, n / 2048}));
    std::vector<double> m((size_t)nth, 0.0);
    if (nth == 1) {
        fill(0, n, m[0]);
        return m[0];
    }
    std::vector<std::thread> pool;
    pool.reserve((size_t)nth - 1);
    for (int64_t i = 1; i < nth; ++i) {
        try {
            pool.emplace_back([&, i] { fill(n * i / nth, n * (i + 1) / nth, m[(size_t)i]); });
        } catch (const std::system_error &) {
            fill(n * i / nth, n * (i + 1) / nth, m[(size_t)i]);  // no thread available: run it here
        }
    }
    fill(0, n / nth, m[0]);
    for (auto &th : pool) th.join();
    return *std::max_element(m.begin(), m.end());
}

void TableBuilder::plan(const int64_t *pf, const int64_t *pl, const uint8_t *rf, int64_t np) {
    segs_.clear();
    ts_ = TableSet{};
    const int NMH = ps_.nmh, NHO = ps_.nho;
    int64_t row = 0, b = 0, tiles = 0, nab = 0, nprec = 0;
    kr_bound_ = 0;
    while (b < np) {
        Segment s;
        if (rf[b]) {
            s.rf = true;
            s.first = pf[b];
            s.last = pl[b];
            ++b;
        } else {
            int64_t e = b;
            while (e + 1 < np && !rf[e + 1]) ++e;
            s.first = pf[b];
            s.last = pl[e];
            b = e + 1;
        }
        for (int64_t n = s.first + 1; n <= s.last; ++n) s.nadc += sq_.adc[(size_t)n] ? 1 : 0;
        s.row0 = row;
        row += s.nadc;
        if (!s.rf) {
            s.tiles = (s.nadc + 15) / 16;
            s.nsb = (s.tiles + kBlockTiles - 1) / kBlockTiles;
            s.nab = (s.tiles + kAnchorTiles - 1) / kAnchorTiles;
            s.kref_off = tiles * kRefsPerTile * 4;
            s.bho_off = tiles * NMH * 64;
            s.s_off = tiles * 16;
            s.kend_off = nprec * (tl_.n + 1);
            s.anc_off = nab * NHO;
            tiles += s.tiles;
            nab += s.nab;
            ++nprec;
            kr_bound_ += s.tiles * NI_ * 64;
        }
        segs_.push_back(s);
    }
    // Every table sized once (growing them segment by segment copies
    // everything accumulated so far at each reallocation).
    ts_.kref.assign((size_t)(tiles * kRefsPerTile * 4), 0.0);
    ts_.bho.assign((size_t)(tiles * NMH * 64), half8_t{});
    ts_.tk.assign((size_t)(tiles * 16), 0.0f);
    ts_.kr0.assign((size_t)(tiles * 16), 0.0);
    ts_.anc.assign((size_t)(nab * NHO), 0.0);
    ts_.kend.assign((size_t)(nprec * (tl_.n + 1)), 0.0);
}

void TableBuilder::fill(size_t i) {
    Segment &s = segs_[i];
    if (s.rf) return;
    const SeqHost &sq = sq_;
    const TermList &tl = tl_;
    const PhaseSplit &ps = ps_;
    const int NI = NI_, flags = flags_;
    const int NMH = ps.nmh, NHO = ps.nho, N1 = ps.n1;
    const int64_t nn = sq.nn;
    const bool ho1 = flags & HOM_HO1, ho0 = flags & HOM_HO0;
    auto wave_of = [&](int code) -> const double * {
        if (code == kTermDW || code == 0) return nullptr;
        if (code <= 3 && !ho1) return code == 1 ? sq.gx : code == 2 ? sq.gy : sq.gz;
        return sq.coef + (size_t)code * nn;
    };
    std::vector<const double *> wave(tl.n);
    for (int q = 0; q < tl.n; ++q) wave[q] = wave_of(tl.code[q]);
    // positions of the FP64, first-order and higher-order terms inside the
    // full list
    int pos64[4] = {-1, -1, -1, -1}, pos1[4] = {-1, -1, -1, -1}, posho[12];
    for (int q = 0; q < tl.n; ++q) {
        for (int j = 0; j < 4; ++j)
            if (ps.code64[j] && tl.code[q] == ps.code64[j]) pos64[j] = q;
        for (int j = 0; j < N1; ++j)
            if (tl.code[q] == ps.code1[j]) pos1[j] = q;
        for (int j = 0; j < NHO; ++j)
            if (tl.code[q] == ps.codeho[j]) posho[j] = q;
    }
    const double *h0 = sq.coef;
    const double g = sq.gamma;
    // B-side scale of each higher-order term: a power of two, so the product
    // equals ldexp(., kHOOut - aexp) exactly
    double bscale[12], bscale1[4];
    for (int q = 0; q < NHO; ++q) bscale[q] = std::ldexp(1.0, kHOOut - ps.aexp[q]);
    for (int q = 0; q < N1; ++q) bscale1[q] = std::ldexp(1.0, kHOOut - ps.aexp1[q]);
    auto &kr = ts_.kr, &kref = ts_.kref, &kr0 = ts_.kr0, &kend = ts_.kend, &anc = ts_.anc;
    auto &bho = ts_.bho;
    auto &tk = ts_.tk;
    // Phase 1 (sequential): the trapezoid integrals at every acquired
    // node (Kr of each term, the time, the zeroth order).
    const int nt = tl.n;
    std::vector<double> KrS((size_t)(s.nadc * nt)), tS((size_t)s.nadc), i0S((size_t)s.nadc);
    std::vector<double> I(nt, 0.0);
    double I0 = 0.0, t = 0.0;
    int64_t k = 0;
    for (int64_t n = s.first + 1; n <= s.last; ++n) {
        const double dt = sq.dt[(size_t)(n - 1)];
        for (int q = 0; q < nt; ++q)
            if (wave[q]) I[q] += (wave[q][n - 1] + wave[q][n]) * (dt / 2.0);
        if (ho0) I0 += (h0[n - 1] + h0[n]) * (dt / 2.0);
        t += dt;
        if (!sq.adc[(size_t)n]) continue;
        double *kr_k = &KrS[(size_t)(k * nt)];
        for (int q = 0; q < nt; ++q) kr_k[q] = tl.code[q] == kTermDW ? -t / kTwoPi : -g * I[q];
        tS[(size_t)k] = t;
        i0S[(size_t)k] = I0;
        ++k;
    }
    // higher-order anchors: the values at the first sample of each 256-sample block
    for (int64_t ab = 0; ab < s.nab; ++ab)
        for (int q = 0; q < NHO; ++q)
            anc[(size_t)(s.anc_off + ab * NHO + q)] = KrS[(size_t)(ab * 16 * kAnchorTiles * nt + posho[q])];
    // first-order anchors: the values at each tile's reference sample
    for (int64_t tau = 0; tau < s.tiles; ++tau)
        for (int h = 0; h < kRefsPerTile; ++h) {
            const int64_t r = ref_sample(tau, h, s.nadc);
            for (int q = 0; q < 4; ++q)
                kref[(size_t)(s.kref_off + (tau * kRefsPerTile + h) * 4 + q)] =
                    pos64[q] >= 0 ? KrS[(size_t)(r * nt + pos64[q])] : 0.0;
        }
    // Phase 2 (samples independent, split over threads): FP64 and
    // split-FP16 table entries, and the largest scaled offset.
    _Float16 *const bho16 = reinterpret_cast<_Float16 *>(bho.data());
    // The largest |B| is reported relative to its bound: higher orders
    // |B| / kHOBMax, first order |B| / kFOBMax (the segment falls back to
    // the all-FP64 body above 1).
    auto fill_range = [&](int64_t k0, int64_t k1, double &bm_out) {
        double bm = 0.0;  // local: the threads' results share a cache line
        for (int64_t kk = k0; kk < k1; ++kk) {
            const double *kr_k = &KrS[(size_t)(kk * nt)];
            const int64_t tau = kk / 16, j = kk % 16, ab = kk / (16 * kAnchorTiles);
            const double *kr_ref = &KrS[(size_t)(ref_sample(tau, (int)(j / kRefHalf), s.nadc) * nt)];
            // first-order slot 3 NHO + p N1 + term holds part fo_bpart(p)
            // of the three-way split offset from the tile reference
            for (int term = 0; term < N1; ++term) {
                const double bval = (kr_k[pos1[term]] - kr_ref[pos1[term]]) * bscale1[term];
                bm = std::max(bm, std::fabs(bval) / kFOBMax);
                _Float16 parts[3];
                host_split16x3(bval, parts);
                for (int prod = 0; prod < kFOParts; ++prod) {
                    const int slot = 3 * NHO + prod * N1 + term, m = slot / 32, sl = slot % 32;
                    bho16[(size_t)(s.bho_off + (tau * NMH + m) * 64 + 16 * (sl / 8) + j) * 8 + sl % 8] =
                        parts[fo_bpart(prod)];
                }
            }
            // slots term, NHO + term, 2 NHO + term hold the B-side hi, hi,
            // lo of the term's split offset (the A side holds hi, lo, hi)
            for (int term = 0; term < NHO; ++term) {
                const double bval =
                    (kr_k[posho[term]] - anc[(size_t)(s.anc_off + ab * NHO + term)]) * bscale[term];
                bm = std::max(bm, std::fabs(bval) / kHOBMax);
                _Float16 hi, lo;
                host_split16(bval, hi, lo);
                for (int kind = 0; kind < 3; ++kind) {
                    const int slot = kind * NHO + term, m = slot / 32, sl = slot % 32;
                    // element sl % 8 of the 8-half chunk, stored as a scalar
                    bho16[(size_t)(s.bho_off + (tau * NMH + m) * 64 + 16 * (sl / 8) + j) * 8 + sl % 8] =
                        kind == 2 ? lo : hi;
                }
            }
            tk[(size_t)(s.s_off + kk)] = (float)tS[(size_t)kk];
            kr0[(size_t)(s.s_off + kk)] = -g * i0S[(size_t)kk];
        }
        bm_out = bm;
    };
    const double bmax = parallel_max(s.nadc, fill_range);
    s.fallback = !(bmax <= 1.0);
    if (s.fallback) {
        // All-FP64 table of every term, only for segments that need it.
        s.kr_off = (int64_t)kr.size();
        kr.resize(kr.size() + (size_t)(s.tiles * NI * 64), 0.0);
        for (int64_t kk = 0; kk < s.nadc; ++kk) {
            const int64_t tau = kk / 16, j = kk % 16;
            for (int q = 0; q < nt; ++q)
                kr[(size_t)(s.kr_off + (tau * NI + q / 4) * 64 + 16 * (q % 4) + j)] =
                    KrS[(size_t)(kk * nt + q)];
        }
    }
    for (int q = 0; q < tl.n; ++q)
        kend[(size_t)(s.kend_off + q)] = tl.code[q] == kTermDW ? -t / kTwoPi : -g * I[q];
    kend[(size_t)(s.kend_off + tl.n)] = -g * I0;
    s.T = t;
}

void build_tables(const SeqHost &sq, const TermList &tl, const PhaseSplit &ps, int NI, int flags,
                  const int64_t *pf, const int64_t *pl, const uint8_t *rf, int64_t np,
                  std::vector<Segment> &segs, TableSet &ts) {
    TableBuilder tb(sq, tl, ps, NI, flags, segs, ts);
    tb.plan(pf, pl, rf, np);
    for (size_t i = 0; i < segs.size(); ++i) tb.fill(i);
}

void mixed_free(hom_plan *p) {
    if (!p || !p->mixed) return;
    MixedWorkspace *W = p->mixed;
    (void)hipStreamSynchronize(p->stream);  // releases proceed regardless
    free_dev(W->basisA); free_dev(W->basis64); free_dev(W->aho); free_dev(W->cdec);
    free_dev(W->bop); free_dev(W->csc); free_dev(W->th0); free_dev(W->lm); free_dev(W->m0max);
    free_dev(W->seed); free_dev(W->kr); free_dev(W->kref); free_dev(W->bho); free_dev(W->anc);
    free_dev(W->tk); free_dev(W->kr0); free_dev(W->kend); free_dev(W->partials);
    for (int b = 0; b < 2; ++b) {
        if (W->prep_done[b]) (void)hipEventDestroy(W->prep_done[b]);
        if (W->main_done[b]) (void)hipEventDestroy(W->main_done[b]);
    }
    if (W->joined) (void)hipEventDestroy(W->joined);
    if (W->main_stream) (void)hipStreamDestroy(W->main_stream);
    delete W;
    p->mixed = nullptr;
}

namespace {
// Device copy of one precession segment's table entries; a segment that
// falls back also copies its all-FP64 table, into a buffer sized on the
// first fallback for every segment's (kr_bound entries).
int upload_segment_tables(hom_plan *p, MixedWorkspace &W, const TableSet &ts, const Segment &s,
                          int64_t kr_bound) {
    const int64_t nho = W.split.nho, nmh = W.split.nmh;
    int rc = upload_range(p, W.kref, ts.kref, s.kref_off, s.tiles * kRefsPerTile * 4);
    rc = rc ? rc : upload_range(p, W.bho, ts.bho, s.bho_off, s.tiles * nmh * 64);
    rc = rc ? rc : upload_range(p, W.tk, ts.tk, s.s_off, s.tiles * 16);
    rc = rc ? rc : upload_range(p, W.kr0, ts.kr0, s.s_off, s.tiles * 16);
    rc = rc ? rc : upload_range(p, W.anc, ts.anc, s.anc_off, s.nab * nho);
    rc = rc ? rc : upload_range(p, W.kend, ts.kend, s.kend_off, (int64_t)W.terms.n + 1);
    if (rc || !s.fallback) return rc;
    if (!W.kr && (rc = dev_alloc(p, &W.kr, (size_t)kr_bound))) return rc;
    return upload_range(p, W.kr, ts.kr, s.kr_off, s.tiles * W.NI * 64);
}

// Spin tile of one segment of nsb sample blocks: the fixed tile, or the
// largest power of two up to the bound whose work items fill the device's
// workgroup slots (two per CU) in whole waves best.  A smaller tile is taken
// only when it fills the waves more than 5 % better, relative (each halving
// doubles the partials the finalize sums).  Measured on MI355X (main + finalize):
// 10,000 spins x 1 coil x 5000 samples 0.366 ms at 256 groups, 0.095 at the
// 32 this picks; 350,000 spins x 8 coils x 4 x 4750: 4.19 ms at 256, 3.90 at 64.
int segment_spin_tile(const hom_plan *p, int64_t ngroups, int64_t nsb) {
    if (p->spin_tile_fixed) return p->spin_tile;
    const int64_t slots = 2 * (int64_t)std::max(p->cu_count, 1);
    auto fill = [&](int g) {
        const int64_t items = (ngroups + g - 1) / g * nsb;
        const int64_t waves = std::max<int64_t>((items + slots - 1) / slots, 1);
        return (double)items / (double)(waves * slots);
    };
    int best = p->spin_tile;
    double best_fill = fill(best);
    for (int g = p->spin_tile / 2; g >= 16; g /= 2) {
        const double f = fill(g);
        if (f > best_fill * 1.05) {
            best = g;
            best_fill = f;
        }
    }
    return best;
}

// On destruction, when the second stream was used: the first stream waits
// for everything queued so far on the second (or, if that cannot be
// queued, the host waits for the second stream).
struct StreamJoin {
    hipStream_t first, second;
    hipEvent_t ev;
    bool used = false;
    ~StreamJoin() {
        if (!used) return;
        if (hipEventRecord(ev, second) != hipSuccess || hipStreamWaitEvent(first, ev, 0) != hipSuccess)
            (void)hipStreamSynchronize(second);
    }
};

// The block loop of the mixed path over parts [pf, pl] (rf marks excitation
// parts), writing sum(ADC) x Nc complex double rows into sig (column stride ld).
int mixed_exec(hom_plan *p, const int64_t *pf, const int64_t *pl, const uint8_t *rf, int64_t np,
               int flags, double2 *sig, int64_t ld) {
    if (!p->mixed) p->mixed = new MixedWorkspace();
    MixedWorkspace &W = *p->mixed;
    p->mixed_segments = p->mixed_fp64_segments = 0;
    const int64_t ns = p->ns, nc = p->nc;
    int rc = HOM_OK;
    const int mflags = flags & (HOM_HO0 | HOM_HO1 | HOM_HO2 | HOM_HO3 | HOM_DW_PRECESSION);
    const int NCB = (int)((2 * nc + 15) / 16);
    const int64_t ngroups = (ns + 15) / 16;
    const double *dx = (const double *)p->d.x, *dy = (const double *)p->d.y,
                 *dz = (const double *)p->d.z, *ddw = (const double *)p->d.dw;

    // Spin-side operands that depend only on the phantom, the CSM and the mask.
    if (W.basis_flags != mflags || W.ngroups != ngroups || W.NCB != NCB) {
        if (const int e = check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize")) return e;
        W.terms = make_terms(mflags);
        W.split = make_split(mflags, p->h_basis_max);
        W.NI = (W.terms.n + 3) / 4;
        W.NCB = NCB;
        W.ngroups = ngroups;
        free_dev(W.basisA); free_dev(W.basis64); free_dev(W.aho); free_dev(W.cdec);
        free_dev(W.bop); free_dev(W.csc); free_dev(W.th0); free_dev(W.lm); free_dev(W.m0max);
        const int nmh = std::max(W.split.nmh, 1);
        if (hom::dev_malloc((void **)&W.basisA, sizeof(double) * (size_t)(ngroups * W.NI * 64)) != hipSuccess ||
            hom::dev_malloc((void **)&W.basis64, sizeof(double) * (size_t)((ngroups + 1) * 64)) != hipSuccess ||
            hom::dev_malloc((void **)&W.aho, sizeof(half8_t) * (size_t)(ngroups * nmh * 64)) != hipSuccess ||
            hom::dev_malloc((void **)&W.cdec, sizeof(float) * (size_t)(ngroups * 16)) != hipSuccess ||
            hom::dev_malloc((void **)&W.bop, sizeof(half8_t) * (size_t)(ngroups * NCB * 2 * 64)) != hipSuccess ||
            hom::dev_malloc((void **)&W.csc, sizeof(double) * (size_t)(8 * NCB)) != hipSuccess ||
            hom::dev_malloc((void **)&W.th0, sizeof(double) * (size_t)(2 * ngroups * 16)) != hipSuccess ||
            hom::dev_malloc((void **)&W.lm, sizeof(float) * (size_t)(2 * ngroups * 16)) != hipSuccess ||
            hom::dev_malloc((void **)&W.m0max, sizeof(float) * 2) != hipSuccess) {
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
                           p->stream, ns, nc, ngroups, NCB, (const double2 *)p->d.csm, ns,
                           (const double *)W.csc, W.bop);
        hipLaunchKernelGGL(k_mix_basis, dim3(blocks_for(ngroups * W.NI * 64, 256)), dim3(256), 0,
                           p->stream, ns, ngroups, W.NI, W.terms, dx, dy, dz, ddw, W.basisA);
        hipLaunchKernelGGL(k_mix_basis64, dim3(blocks_for((ngroups + 1) * 64, 256)), dim3(256), 0,
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
    // When they are not cached, the tables are planned here (segments, sizes,
    // device buffers) and each segment's entries are computed and uploaded
    // just before its launches, so the host fills segment i + 1 while the
    // device runs segment i (the whole build used to precede the first
    // launch: 21 ms on the host at 1M spins x 16 x 8000 samples).
    SeqHost sq;
    sq.nn = p->nn;
    sq.gamma = p->gamma;
    sq.dt = p->h_dt.data();
    sq.gx = p->h_gx.data();
    sq.gy = p->h_gy.data();
    sq.gz = p->h_gz.data();
    sq.coef = p->h_coef.data();
    sq.adc = p->h_adc.data();
    TableSet ts;
    TableBuilder tb(sq, W.terms, W.split, W.NI, mflags, W.segs, ts);
    const bool building = W.table_flags != mflags || W.table_parts != key;
    if (building) {
        if (const int e = check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize")) return e;
        W.table_flags = -1;  // valid again once every segment is uploaded
        tb.plan(pf, pl, rf, np);
        rc = dev_alloc(p, &W.kref, ts.kref.size());
        rc = rc ? rc : dev_alloc(p, &W.bho, ts.bho.size());
        rc = rc ? rc : dev_alloc(p, &W.anc, ts.anc.size());
        rc = rc ? rc : dev_alloc(p, &W.tk, ts.tk.size());
        rc = rc ? rc : dev_alloc(p, &W.kr0, ts.kr0.size());
        rc = rc ? rc : dev_alloc(p, &W.kend, ts.kend.size());
        free_dev(W.kr);  // allocated when the first segment falls back
        if (rc) return rc;
    }
    // Spin tiles (16-spin groups per work item) of the segments; partial and
    // seed buffers sized for the largest segment.
    auto stiles_of = [&](int g) {
        const int64_t stile_spins = (int64_t)g * 16;
        return std::max<int64_t>((ns + stile_spins - 1) / stile_spins, 1);
    };
    std::vector<int> seg_gpt(W.segs.size(), p->spin_tile);
    int64_t need = 0, need_seed = 0, longest = -1;
    p->last_spin_tile = 0;
    for (size_t i = 0; i < W.segs.size(); ++i) {
        const Segment &s = W.segs[i];
        if (s.rf) continue;
        seg_gpt[i] = segment_spin_tile(p, ngroups, s.nsb);
        need = std::max<int64_t>(need, stiles_of(seg_gpt[i]) * s.tiles * 16 * 16 * NCB);
        need_seed = std::max<int64_t>(need_seed, s.nab * ngroups * 16 + 16);
        if (s.nadc > longest) {
            longest = s.nadc;
            p->last_spin_tile = seg_gpt[i];
        }
    }
    if (need > W.partial_cap || need_seed > W.seed_cap) {
        if (const int e = check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize")) return e;
        if (need > W.partial_cap) {
            free_dev(W.partials);
            W.partial_cap = 0;
            if (hom::dev_malloc((void **)&W.partials, sizeof(float) * (size_t)(2 * need)) != hipSuccess) {
                W.partials = nullptr;
                return set_error(p, HOM_ERR_MEMORY, "device allocation failed for mixed partials");
            }
            W.partial_cap = need;
        }
        if (need_seed > W.seed_cap) {
            free_dev(W.seed);
            W.seed_cap = 0;
            if (hom::dev_malloc((void **)&W.seed, sizeof(double) * (size_t)(2 * need_seed)) != hipSuccess) {
                W.seed = nullptr;
                return set_error(p, HOM_ERR_MEMORY, "device allocation failed for phase seeds");
            }
            W.seed_cap = need_seed;
        }
    }

    // Two streams.  The main kernels run on W.main_stream; the RF blocks and
    // each segment's prep, finalize and end-state kernels run on the plan's
    // stream, so they fill the device beside a main kernel instead of
    // between two.  Segment k with samples uses slot k % 2 of th0, lm,
    // m0max, seed and partials.  The plan's stream runs prep(k),
    // finalize(k - 1) (after main(k - 1)), end(k), prep(k + 1), ...: prep(k + 1)
    // overlaps main(k); slot k % 2 is rewritten by prep(k + 2) only after
    // finalize(k), which waited for main(k); main(k + 2) waits for prep(k + 2)
    // and so for finalize(k), the last reader of its partials.  Every exit
    // makes the plan's stream wait for the main stream, so synchronising the
    // plan's stream covers both.
    if (!W.main_stream) {
        hipStream_t st = nullptr;
        hipEvent_t ev[5] = {};
        bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
        for (int i = 0; ok && i < 5; ++i) ok = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) == hipSuccess;
        if (!ok) {
            for (hipEvent_t e : ev)
                if (e) (void)hipEventDestroy(e);
            if (st) (void)hipStreamDestroy(st);
            (void)hipGetLastError();
            return set_error(p, HOM_ERR_DEVICE, "stream or event creation failed for the mixed path");
        }
        W.main_stream = st;
        W.prep_done[0] = ev[0];
        W.prep_done[1] = ev[1];
        W.main_done[0] = ev[2];
        W.main_done[1] = ev[3];
        W.joined = ev[4];
    }
    StreamJoin join{p->stream, W.main_stream, W.joined};
    struct Fin {
        bool on;
        int slot;
        int64_t nadc, nstiles, rows, s_off, row0;
    };
    Fin fin{};
    int64_t nsamp_segs = 0;
    auto finalize = [&](Fin &f) -> int {
        if (!f.on) return HOM_OK;
        f.on = false;
        int r = check_hip(p, hipStreamWaitEvent(p->stream, W.main_done[f.slot], 0), "hipStreamWaitEvent");
        if (r) return r;
        {
            ScopedTimer timer(p, HOM_K_MIXED_FINALIZE);
            hipLaunchKernelGGL(k_mix_finalize, dim3(blocks_for(f.nadc * nc, 256)), dim3(256), 0, p->stream,
                               f.nadc, nc, f.nstiles, f.rows, (int64_t)(16 * NCB),
                               (const float *)(W.partials + f.slot * W.partial_cap),
                               (const float *)(W.m0max + f.slot), (const double *)W.csc,
                               (const double *)(W.kr0 + f.s_off), sig + f.row0, ld);
        }
        return check_hip(p, hipGetLastError(), "mixed finalize");
    };
    for (size_t si = 0; si < W.segs.size(); ++si) {
        if (building && !W.segs[si].rf) {
            tb.fill(si);
            if ((rc = upload_segment_tables(p, W, ts, W.segs[si], tb.kr_bound()))) return rc;
        }
        const Segment &s = W.segs[si];
        if (s.rf) {
            // RF blocks: f64 per-block excitation (+ reduction of RF ADC samples).
            if (s.nadc > p->max_adc)
                return set_error(p, HOM_ERR_DIMENSION, "The block acquires more samples than max_adc.");
            if ((rc = ensure_state_cols(p, s.nadc))) return rc;
            rc = check_hip(p, launch_plain_excitation(p, s.first, s.last, flags), "excitation kernel");
            if (rc) return rc;
            if (s.nadc) {
                const int64_t need_p = coil_partial_elements(ns, s.nadc, nc);
                if (need_p > p->d.partial_capacity) {
                    if (const int e = check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize")) return e;
                    if (p->d.partials) (void)hipFree(p->d.partials);
                    p->d.partials = nullptr;
                    p->d.partial_capacity = 0;
                    if (hom::dev_malloc(&p->d.partials, (size_t)need_p * p->complex_size()) != hipSuccess)
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
            ++p->mixed_segments;
            if (s.fallback) ++p->mixed_fp64_segments;
            const int slot = (int)(nsamp_segs++ & 1);
            double *th0 = W.th0 + slot * ngroups * 16;
            float *lm = W.lm + slot * ngroups * 16;
            float *m0max = W.m0max + slot;
            double *seed = W.seed + slot * W.seed_cap;
            float *partials = W.partials + slot * W.partial_cap;
            {
                ScopedTimer timer(p, HOM_K_MIXED_PREP);
                hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, p->stream, m0max, (int64_t)1, 0.0f);
                hipLaunchKernelGGL(k_mix_m0max, dim3((unsigned)std::min<int64_t>(blocks_for(ns, 256), 256)),
                                   dim3(256), 0, p->stream, ns,
                                   (const double2 *)p->d.mxy, m0max);
                hipLaunchKernelGGL(k_mix_seg, dim3(blocks_for(ngroups * 16, 256)), dim3(256), 0,
                                   p->stream, ns, ngroups, (const double2 *)p->d.mxy,
                                   (const float *)m0max, th0, lm);
                if (!s.fallback)
                    hipLaunchKernelGGL(k_mix_seed, dim3(blocks_for(ngroups * 16, 256)), dim3(256),
                                       0, p->stream, ns, ngroups, s.nab, W.split, dx, dy, dz, ddw,
                                       (const double *)th0, (const double *)(W.anc + s.anc_off),
                                       seed);
            }
            rc = check_hip(p, hipGetLastError(), "mixed prep");
            rc = rc ? rc : check_hip(p, hipEventRecord(W.prep_done[slot], p->stream), "hipEventRecord");
            rc = rc ? rc : check_hip(p, hipStreamWaitEvent(W.main_stream, W.prep_done[slot], 0),
                                     "hipStreamWaitEvent");
            if (rc) return rc;
            join.used = true;
            const int gpt = seg_gpt[si];
            const int64_t nstiles = stiles_of(gpt);
            const int64_t nwork = s.nsb * nstiles;
            for (int cb0 = 0; cb0 < NCB; cb0 += 4) {
                const int ncb = std::min(4, NCB - cb0);
                MainOps o{};
                o.ngroups = ngroups;
                o.gpt = gpt;
                o.cb0 = cb0;
                o.ncb_total = NCB;
                o.tile_lo = 0;
                o.tile_hi = s.tiles;
                o.sb_lo = 0;
                o.ab_lo = 0;
                o.basis64 = W.basis64;
                o.aho = W.aho;
                o.cdec = W.cdec;
                o.lm = lm;
                o.seed = seed;
                o.bop = W.bop;
                o.kref = W.kref + s.kref_off;
                o.bho = W.bho + s.bho_off;
                o.tk = W.tk + s.s_off;
                o.partials = partials;
                o.basisA = W.basisA;
                o.th0 = th0;
                o.kr = s.fallback ? W.kr + s.kr_off : nullptr;
                ScopedTimer timer(p, HOM_K_MIXED_SEGMENT, W.main_stream);
                rc = check_hip(p, launch_main(ncb, s.fallback, W.split.nmh, W.NI, W.main_stream, o, s.nsb,
                                              nwork),
                               "mixed main kernel");
                if (rc) return rc;
            }
            rc = check_hip(p, hipEventRecord(W.main_done[slot], W.main_stream), "hipEventRecord");
            if (rc) return rc;
            // the previous segment's finalize; this one's waits for the next
            // segment's prep (or the end of the loop)
            if ((rc = finalize(fin))) return rc;
            fin = Fin{true, slot, s.nadc, nstiles, s.tiles * 16, s.s_off, s.row0};
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
    if ((rc = finalize(fin))) return rc;
    if (building) {
        W.table_flags = mflags;
        W.table_parts = key;
    }
    return HOM_OK;
}
}  // namespace

int mixed_run(hom_plan *p, const int64_t *pf, const int64_t *pl, const uint8_t *rf, int64_t np,
              int flags) {
    return mixed_exec(p, pf, pl, rf, np, flags, (double2 *)p->signal, std::max<int64_t>(p->signal_rows, 1));
}

int mixed_precession_block(hom_plan *p, int64_t first, int64_t last, int flags, void *out, int64_t ld) {
    const uint8_t rf0 = 0;
    return mixed_exec(p, &first, &last, &rf0, 1, flags, (double2 *)out, std::max<int64_t>(ld, 1));
}

}  // namespace hom
