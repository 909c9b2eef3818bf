// serve.hip -- continuous-batching simulation server (include/hom_serve.h).
//
// Device side: persistent batched versions of the mixed path's per-segment
// kernels (mixed_common.h bodies).  Each reads one step descriptor from
// device memory: per-request records (JobDesc), per-kernel request lists
// with block prefix sums, and the fused-kernel work entries, each tagged with
// its kernel class (split phase with NMH offset MFMAs or all-FP64 phase with
// NI FP64 MFMAs, times NCB coil blocks).  One fused kernel serves every
// class, switching per work item, so the step is a fixed chain of seven
// kernels with fixed grids: one captured hipGraph replays every step and only
// the descriptor changes.
//
// Fused-kernel items are spread over the eight XCDs (block b runs on XCD
// b % 8): XCD x takes items [x T/8, (x+1) T/8) from its own atomic counter,
// so consecutive items, which share a spin tile, hit that XCD's L2.  Every
// workgroup leaves its loop once its counter passes the end of its range.
//
// Host side: resident store (phantoms, CSMs, sequences in HBM), a first-fit
// device arena that requests are admitted against, a pinned host arena for
// results, and the step scheduler (admission in arrival order, one chunk of
// whole anchor blocks per active request per step, RF blocks of a request
// launched before the step's graph, retirement after the step).
#include "mixed_common.h"

#include "../../include/hom_serve.h"

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

namespace hom {
namespace serve {

constexpr int kMaxJobs = HOM_SERVE_MAX_ACTIVE;
// run_class cases 0..7: split phase, nmh 1..2 x ncb 1..4; 8..23: all-FP64,
// NI 1..4 x ncb 1..4
constexpr int kSplitClasses = 8;
constexpr int kMaxMain = 4 * kMaxJobs;
constexpr int kLightGrid = 1024;
constexpr int64_t kAlign = 256;

struct JobDesc {
    int64_t ns, ngroups, nc, nstiles;
    const double *x, *y, *z, *dw, *rho, *T1, *T2;
    double2 *mxy;
    double *mz;
    double *th0;
    float *lm;
    float *m0max;
    const double *csc;
    double *seed;
    const double *anc;
    int64_t nab;
    int64_t nadc, rows, cols;  // chunk rows of the partials, signal columns of them
    const float *partials;
    const double *kr0;
    double2 *sig;
    int64_t ld;
    const double *kend;
    double T;
    int32_t do_seg, do_seed;  // segment start (th0, lm, scale) / anchor seeds this step
    PhaseSplit ps;
    TermList tl;
};

// Requests of one light kernel: entry e covers blocks [prefix[e], prefix[e+1]).
struct ListDesc {
    int32_t n;
    int32_t job[kMaxJobs];
    int64_t prefix[kMaxJobs + 1];
};

struct MainDesc {
    MainOps o;
    int64_t nsb;      // sample blocks of the chunk
    int64_t item_lo;  // first item of the entry
    int32_t cls;      // kernel class
};

struct StepDesc {
    unsigned int ctr[8];       // per-XCD item counters (zero on upload)
    float m0acc[kMaxJobs];     // per job slot: max |M0| of a starting segment (zero on upload)
    int32_t n_main;
    int64_t total_items;
    ListDesc m0, ss, fin, end;  // m0max; seg + seed; finalize; end of segment
    JobDesc jobs[kMaxJobs];
    MainDesc main[kMaxMain];
};

// ------------------------------------------------------------- kernels --

__device__ inline int find_entry(const ListDesc &L, int64_t b) {
    int lo = 0, hi = L.n;  // prefix[lo] <= b < prefix[hi]
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (L.prefix[m] <= b) lo = m;
        else hi = m;
    }
    return lo;
}

// Max |M0| of every request starting a segment, into m0acc[slot].
__global__ __launch_bounds__(256) void k_srv_m0max(const StepDesc *__restrict__ sd,
                                                   float *__restrict__ m0acc) {
    const ListDesc &L = sd->m0;
    const int64_t total = L.prefix[L.n];
    for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
        const int e = find_entry(L, b);
        const JobDesc &J = sd->jobs[L.job[e]];
        m0max_block((b - L.prefix[e]) * 256 + threadIdx.x, J.ns, J.mxy, m0acc + L.job[e]);
    }
}

// Per spin slot: segment start (th0, lm; the scale is kept in the request's
// memory for the later chunks) and then the chunk's anchor seeds, which read
// the th0 the same thread just wrote.
__global__ __launch_bounds__(256) void k_srv_segseed(const StepDesc *__restrict__ sd,
                                                     const float *__restrict__ m0acc) {
    const ListDesc &L = sd->ss;
    const int64_t total = L.prefix[L.n];
    for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
        const int e = find_entry(L, b);
        const int slot = L.job[e];
        const JobDesc &J = sd->jobs[slot];
        const int64_t gi = (b - L.prefix[e]) * 256 + threadIdx.x;
        if (J.do_seg) {
            const float m0 = m0acc[slot];
            if (gi == 0) *J.m0max = m0;
            if (gi < J.ngroups * 16) seg_elem(gi, J.ns, J.mxy, m0, J.th0, J.lm);
        }
        if (J.do_seed && gi < J.ngroups * 16)
            seed_elem(gi, J.ns, J.ngroups, J.nab, J.ps, J.x, J.y, J.z, J.dw, J.th0, J.anc, J.seed);
    }
}

// Finalize of every chunk (blocks [0, F)) and end of every finished segment
// (blocks [F, F + E)); the two touch disjoint data.
__global__ __launch_bounds__(256) void k_srv_finend(const StepDesc *__restrict__ sd) {
    const ListDesc &F = sd->fin, &E = sd->end;
    const int64_t nf = F.prefix[F.n], total = nf + E.prefix[E.n];
    for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
        if (b < nf) {
            const int e = find_entry(F, b);
            const JobDesc &J = sd->jobs[F.job[e]];
            const int64_t idx = (b - F.prefix[e]) * 256 + threadIdx.x;
            if (idx < J.nadc * J.nc)
                finalize_elem(idx, J.nc, J.nstiles, J.rows, J.cols, J.partials, *J.m0max, J.csc,
                              J.kr0, J.sig, J.ld);
        } else {
            const int e = find_entry(E, b - nf);
            const JobDesc &J = sd->jobs[E.job[e]];
            const int64_t s = (b - nf - E.prefix[e]) * 256 + threadIdx.x;
            if (s < J.ns)
                end_elem(s, J.tl, J.x, J.y, J.z, J.dw, J.rho, J.T1, J.T2, J.kend, J.T, J.mxy, J.mz);
        }
    }
}

__device__ inline int64_t uniform64(int64_t v) {
    const int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Body of one work item for each kernel class.
__device__ __forceinline__ void run_class(int cls, const MainOps &o, int64_t sb, int64_t st,
                                          const SplitSmem &sm) {
    const WorkItem it = make_item(o, sb, st);
    switch (cls) {
#define HOM_SPLIT(C, NMH, NCB) \
    case C: main_split_body<NMH, NCB>(o, it, sm); break;
#define HOM_FP64(C, NI, NCB) \
    case C: main64_body<NI, NCB>(o, it); break;
        HOM_SPLIT(0, 1, 1) HOM_SPLIT(1, 1, 2) HOM_SPLIT(2, 1, 3) HOM_SPLIT(3, 1, 4)
        HOM_SPLIT(4, 2, 1) HOM_SPLIT(5, 2, 2) HOM_SPLIT(6, 2, 3) HOM_SPLIT(7, 2, 4)
        HOM_FP64(8, 1, 1) HOM_FP64(9, 1, 2) HOM_FP64(10, 1, 3) HOM_FP64(11, 1, 4)
        HOM_FP64(12, 2, 1) HOM_FP64(13, 2, 2) HOM_FP64(14, 2, 3) HOM_FP64(15, 2, 4)
        HOM_FP64(16, 3, 1) HOM_FP64(17, 3, 2) HOM_FP64(18, 3, 3) HOM_FP64(19, 3, 4)
        HOM_FP64(20, 4, 1) HOM_FP64(21, 4, 2) HOM_FP64(22, 4, 3) HOM_FP64(23, 4, 4)
#undef HOM_SPLIT
#undef HOM_FP64
        default: break;
    }
}

// The fused kernel of a step: claims items for this workgroup's XCD until its
// range is done; item -> entry by binary search over the entries' item_lo.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HOM_WAVES_PER_EU, HOM_WAVES_PER_EU))) void k_srv_fused(
    const StepDesc *__restrict__ sd, unsigned int *__restrict__ ctr) {
    // LDS of the largest split class (NMH 2, NCB 4) serves every class.
    using L = SplitLds<2, 4>;
    __shared__ __attribute__((aligned(32))) unsigned char rec0[L::kRecBytes];
    __shared__ __attribute__((aligned(32))) unsigned char rec1[L::kRecBytes];
    __shared__ __attribute__((aligned(32))) float anc0[L::kAncBytes / 4];
    __shared__ __attribute__((aligned(32))) float anc1[L::kAncBytes / 4];
    const SplitSmem sm{rec0, rec1, anc0, anc1};
    __shared__ int64_t s_item;
    const int xcd = blockIdx.x % 8;
    const int64_t total = sd->total_items;
    const int64_t lo = total * xcd / 8, hi = total * (xcd + 1) / 8;
    for (;;) {
        if (threadIdx.x == 0) s_item = lo + (int64_t)atomicAdd(ctr + xcd, 1u);
        __syncthreads();
        const int64_t item = uniform64(s_item);
        __syncthreads();
        if (item >= hi) break;
        int b = 0, e = sd->n_main;
        while (e - b > 1) {
            const int m = (b + e) >> 1;
            if (sd->main[m].item_lo <= item) b = m;
            else e = m;
        }
        const MainDesc &md = sd->main[b];
        const int64_t local = item - md.item_lo;
        run_class(md.cls, md.o, local % md.nsb, local / md.nsb, sm);
    }
}

// Request preparation (direct launches at admission).
__global__ void k_srv_prep(int64_t ns, int64_t nc, int64_t ngroups, int NCB, int NI, PhaseSplit ps,
                           TermList tl, const double *x, const double *y, const double *z,
                           const double *dw, const double *T2, const double *rho, const double2 *csm,
                           int64_t ldc, const double *csc, half8_t *bop, double *basis64,
                           half8_t *aho, float *cdec, double *basisA, double2 *mxy, double *mz) {
    const int64_t n_bop = ngroups * NCB * 64, n_b64 = (ngroups + 1) * 64, n_aho = ngroups * ps.nmh * 64,
                  n_cd = ngroups * 16, n_ba = basisA ? ngroups * NI * 64 : 0;
    const int64_t total = n_bop + n_b64 + n_aho + n_cd + n_ba + ns;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t k = i;
        if (k < n_bop) { bop_elem(k, ns, nc, NCB, csm, ldc, csc, bop); continue; }
        k -= n_bop;
        if (k < n_b64) { basis64_elem(k, ns, ps, x, y, z, dw, basis64); continue; }
        k -= n_b64;
        if (k < n_aho) { aho_elem(k, ns, ps, x, y, z, dw, aho); continue; }
        k -= n_aho;
        if (k < n_cd) { cdec_elem(k, ns, T2, cdec); continue; }
        k -= n_cd;
        if (k < n_ba) { basisA_elem(k, ns, NI, tl, x, y, z, dw, basisA); continue; }
        k -= n_ba;
        mxy[k] = double2{0.0, 0.0};  // initial state: Mxy = 0, Mz = rho
        mz[k] = rho[k];
    }
}

int split_class(int nmh, int ncb) { return (nmh - 1) * 4 + (ncb - 1); }
int fp64_class(int NI, int ncb) { return kSplitClasses + (NI - 1) * 4 + (ncb - 1); }

// --------------------------------------------------------------- host --

double now_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int64_t align_up(int64_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one contiguous region (offsets, 256-byte aligned).
struct Arena {
    int64_t size = 0, used = 0, peak = 0;
    std::map<int64_t, int64_t> free_;  // offset -> bytes

    void init(int64_t bytes) {
        size = bytes;
        used = peak = 0;
        free_.clear();
        if (bytes > 0) free_[0] = bytes;
    }
    int64_t alloc(int64_t bytes) {
        bytes = align_up(std::max<int64_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const int64_t off = it->first, rest = it->second - bytes;
            free_.erase(it);
            if (rest > 0) free_[off + bytes] = rest;
            used += bytes;
            peak = std::max(peak, used);
            return off;
        }
        return -1;
    }
    void release(int64_t off, int64_t bytes) {
        bytes = align_up(std::max<int64_t>(bytes, 1));
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// Range maxima over spins of w values per spin (basis terms, coil
// magnitudes), from tables at two block sizes: sub-blocks of kSub spins and
// blocks of kMaxBlock spins.  A query reads at most 2 (kSub - 1) spins from
// the per-spin source, at most 2 (kMaxBlock / kSub - 1) sub-blocks and the
// whole blocks between them.
constexpr int64_t kSub = 64;
constexpr int64_t kMaxBlock = 4096;
constexpr int64_t kSubPerBlock = kMaxBlock / kSub;

struct RangeMax {
    int64_t w = 0;
    std::vector<double> sub, blk;  // (sub-blocks | blocks) x w

    // value(i, k) >= 0: value k of spin i; built column by column (k outer).
    template <typename F>
    void build(int64_t ns, int64_t w_, F value) {
        w = w_;
        const int64_t nsub = (ns + kSub - 1) / kSub, nblk = (ns + kMaxBlock - 1) / kMaxBlock;
        sub.assign((size_t)(nsub * w), 0.0);
        blk.assign((size_t)(nblk * w), 0.0);
        for (int64_t k = 0; k < w; ++k)
            for (int64_t i = 0; i < ns; ++i) {
                double &m = sub[(size_t)((i / kSub) * w + k)];
                m = std::max(m, value(i, k));
            }
        for (int64_t b = 0; b < nsub; ++b)
            for (int64_t k = 0; k < w; ++k) {
                double &m = blk[(size_t)((b / kSubPerBlock) * w + k)];
                m = std::max(m, sub[(size_t)(b * w + k)]);
            }
    }

    // out[k] = max over spins [a, b) of value k, for k < wq <= w.
    template <typename F>
    void query(int64_t a, int64_t b, int64_t wq, F value, double *out) const {
        for (int64_t k = 0; k < wq; ++k) out[k] = 0.0;
        auto direct = [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i)
                for (int64_t k = 0; k < wq; ++k) out[k] = std::max(out[k], value(i, k));
        };
        auto table = [&](const std::vector<double> &t, int64_t lo, int64_t hi) {
            for (int64_t q = lo; q < hi; ++q)
                for (int64_t k = 0; k < wq; ++k) out[k] = std::max(out[k], t[(size_t)(q * w + k)]);
        };
        const int64_t s0 = (a + kSub - 1) / kSub, s1 = b / kSub;  // whole sub-blocks [s0, s1)
        if (s0 >= s1) {
            direct(a, b);
            return;
        }
        direct(a, s0 * kSub);
        direct(s1 * kSub, b);
        const int64_t b0 = (s0 + kSubPerBlock - 1) / kSubPerBlock, b1 = s1 / kSubPerBlock;
        if (b0 >= b1) {
            table(sub, s0, s1);
            return;
        }
        table(sub, s0, b0 * kSubPerBlock);
        table(sub, b1 * kSubPerBlock, s1);
        table(blk, b0, b1);
    }
};

struct ResPhantom {
    int64_t ns = 0;
    double *d = nullptr;  // x, y, z, rho, T1, T2, dw (ns each)
    RangeMax bmax;             // max |basis term j|, j = 0..16
    std::vector<double> hx, hy, hz, hdw;
    int refs = 0;
    bool released = false;
    const double *arr(int i) const { return d + (size_t)i * (size_t)ns; }
};

struct ResCsm {
    int64_t phantom = 0, nc = 0, ns = 0;
    double2 *d = nullptr;          // ns x nc, column stride ns
    RangeMax cmax2;                // max |csm|^2 per coil
    std::vector<double> host;      // interleaved copy for the partial ends of range queries
    int refs = 0;
    bool released = false;
};

struct ResSeq {
    int64_t nn = 0;
    double gamma = 0.0;
    double *d = nullptr;     // gx, gy, gz, df, psi (nn each), b1 (2 nn), dt (nn)
    uint8_t *dadc = nullptr;
    std::vector<double> t, dt, gx, gy, gz, coef;  // coef: 16 x nn
    std::vector<uint8_t> adc, exc;
    int refs = 0;
    bool released = false;
};

struct Job {
    int64_t id = 0;
    hom_request req{};
    int state = HOM_REQ_QUEUED;
    std::string err;
    ResPhantom *ph = nullptr;
    ResCsm *cs = nullptr;
    ResSeq *sq = nullptr;
    int64_t ns = 0, nc = 0, rows = 0, ngroups = 0, nstiles = 0, chunk_sb = 0, max_rf_adc = 0;
    int flags = 0, mflags = 0, NI = 0, NCB = 0, gpt = 256;
    TermList tl{};
    PhaseSplit ps{};
    bool any_fallback = false;
    std::vector<Segment> segs;
    TableSet ts;
    // device layout (offsets into the job's arena block)
    int64_t o_mxy = 0, o_mz = 0, o_b64 = 0, o_aho = 0, o_cdec = 0, o_bop = 0, o_csc = 0, o_th0 = 0,
            o_lm = 0, o_m0 = 0, o_seed = 0, o_ba = -1, o_kr = -1, o_kref = 0, o_bho = 0, o_anc = 0,
            o_tk = 0, o_kr0 = 0, o_kend = 0, o_part = 0, o_sig = 0, o_rfst = -1, o_rfpart = -1;
    int64_t rf_part_cap = 0;
    int64_t dev_bytes = 0, dev_off = -1, host_bytes = 0, host_off = -1;
    char *dev = nullptr;
    double2 *result = nullptr;  // pinned host
    std::vector<double> csc;
    // progress
    size_t seg = 0;
    int64_t sb_next = 0;
    bool finishing = false;
    bool fetching = false;  // a hom_server_result call is copying the signal out
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
    hom_plan view;  // borrowed device pointers for the f64 RF blocks

    template <typename T>
    T *at(int64_t off) const { return off < 0 ? nullptr : reinterpret_cast<T *>(dev + off); }
};

}  // namespace serve
}  // namespace hom

using namespace hom;
using namespace hom::serve;

struct hom_server {
    hom_server_config cfg{};
    hipStream_t stream = nullptr;
    char *dev_base = nullptr;
    char *host_base = nullptr;
    Arena dev_arena, host_arena;
    StepDesc *h_desc = nullptr;  // pinned
    StepDesc *d_desc = nullptr;
    std::map<uint32_t, hipGraphExec_t> graphs;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::string err;
    int64_t next_id = 1;
    std::map<int64_t, std::unique_ptr<ResPhantom>> phantoms;
    std::map<int64_t, std::unique_ptr<ResCsm>> csms;
    std::map<int64_t, std::unique_ptr<ResSeq>> seqs;
    std::map<int64_t, std::unique_ptr<Job>> jobs;
    std::deque<Job *> queue;
    std::vector<Job *> active;
    hom_server_stats stats{};
    std::atomic<int64_t> graph_launches{0}, graphs_captured{0};  // counted outside mu
    std::mutex step_mu;  // one step at a time
    std::thread worker;
    bool stop = false;
    bool fatal = false;
    uint64_t epoch = 0;  // bumped whenever memory or the queue may let a request in
    bool blocked = false;  // worker: the queue head cannot be admitted until epoch changes
    // Failure-handling tests: HOM_SERVE_TEST_FAIL_STEP=n (read at creation)
    // makes step n end in a device error after its kernels have finished
    // normally, so the stop path runs without any fault on the GPU.
    int64_t test_fail_step = 0;
};

namespace {

thread_local std::string g_server_create_error;

int fail(hom_server *s, int code, const std::string &msg) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->err = msg;
    return code;
}

int fail_locked(hom_server *s, int code, const std::string &msg) {
    s->err = msg;
    return code;
}

int hip_fail(hom_server *s, hipError_t e, const char *what) {
    if (e == hipSuccess) return HOM_OK;
    return fail(s, e == hipErrorOutOfMemory ? HOM_ERR_MEMORY : HOM_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

// Max of |basis term j| (j = 0..16) over spins [a, b) of a phantom.
double basis_abs(const ResPhantom &ph, int64_t i, int64_t j) {
    return std::fabs(basis_term((int)j, ph.hx[(size_t)i], ph.hy[(size_t)i], ph.hz[(size_t)i], ph.hdw[(size_t)i]));
}

void range_basis_max(const ResPhantom &ph, int64_t a, int64_t b, double out[17]) {
    ph.bmax.query(a, b, 17, [&](int64_t i, int64_t j) { return basis_abs(ph, i, j); }, out);
}

// |csm[i, c]|^2 of a resident CSM.
double csm_abs2(const ResCsm &cs, int64_t i, int64_t c) {
    const double *v = &cs.host[(size_t)(2 * (i + c * cs.ns))];
    return v[0] * v[0] + v[1] * v[1];
}

// Max |csm| per coil c < nc over spins [a, b).
void range_csm_max(const ResCsm &cs, int64_t a, int64_t b, int64_t nc, std::vector<double> &out) {
    out.assign((size_t)nc, 0.0);
    cs.cmax2.query(a, b, nc, [&](int64_t i, int64_t c) { return csm_abs2(cs, i, c); }, out.data());
    for (double &m : out) m = std::sqrt(m);
}

struct Layout {
    int64_t total = 0;
    int64_t add(int64_t bytes) {
        const int64_t o = total;
        total += align_up(std::max<int64_t>(bytes, 0));
        return o;
    }
};

// Checks a request against the resident store and pins the objects it uses
// (their refcounts keep them alive until the request retires).  Caller holds mu.
int resolve_job_locked(hom_server *s, const hom_request &r, Job &j) {
    auto pit = s->phantoms.find(r.phantom);
    if (pit == s->phantoms.end() || pit->second->released)
        return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown phantom id.");
    auto cit = s->csms.find(r.csm);
    if (cit == s->csms.end() || cit->second->released)
        return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown CSM id.");
    auto qit = s->seqs.find(r.sequence);
    if (qit == s->seqs.end() || qit->second->released)
        return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown sequence id.");
    ResPhantom &ph = *pit->second;
    ResCsm &cs = *cit->second;
    ResSeq &sq = *qit->second;
    if (cs.phantom != r.phantom)
        return fail_locked(s, HOM_ERR_ARGUMENT, "The CSM belongs to another phantom.");
    if (r.spin_count < 1 || r.spin_first < 0 || r.spin_first > ph.ns - r.spin_count)
        return fail_locked(s, HOM_ERR_DIMENSION, "The spin range is outside the phantom.");
    if (r.coil_count < 1 || r.coil_count > cs.nc)
        return fail_locked(s, HOM_ERR_DIMENSION, "The coil count is outside the CSM.");
    if (r.node_first < 0 || r.node_last <= r.node_first || r.node_last >= sq.nn)
        return fail_locked(s, HOM_ERR_DIMENSION, "The node range is outside the sequence.");
    if (r.flags & ~(HOM_HO0 | HOM_HO1 | HOM_HO2 | HOM_HO3 | HOM_DW_EXCITATION | HOM_DW_PRECESSION))
        return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown BlochHighOrder flag bits.");
    j.req = r;
    j.ph = &ph;
    j.cs = &cs;
    j.sq = &sq;
    ph.refs++;
    cs.refs++;
    sq.refs++;
    return HOM_OK;
}

// Fills everything the host can know before admission: segments, tables,
// chunking and the device layout.  Runs without mu (the resident objects are
// pinned and their host copies never change).
int plan_job(hom_server *s, Job &j, int gpt) {
    const hom_request &r = j.req;
    const ResPhantom &ph = *j.ph;
    const ResCsm &cs = *j.cs;
    const ResSeq &sq = *j.sq;
    j.ns = r.spin_count;
    j.nc = r.coil_count;
    j.flags = r.flags;
    j.mflags = r.flags & (HOM_HO0 | HOM_HO1 | HOM_HO2 | HOM_HO3 | HOM_DW_PRECESSION);
    j.gpt = gpt;
    j.NCB = (int)((2 * j.nc + 15) / 16);
    j.ngroups = (j.ns + 15) / 16;
    j.nstiles = std::max<int64_t>((j.ngroups + j.gpt - 1) / j.gpt, 1);
    double bm[17];
    range_basis_max(ph, r.spin_first, r.spin_first + r.spin_count, bm);
    j.tl = make_terms(j.mflags);
    j.ps = make_split(j.mflags, bm);
    j.NI = (j.tl.n + 3) / 4;
    std::vector<double> cmax;
    range_csm_max(cs, r.spin_first, r.spin_first + r.spin_count, j.nc, cmax);
    j.csc.assign((size_t)(8 * j.NCB), 1.0);
    for (int64_t c = 0; c < j.nc; ++c) j.csc[(size_t)c] = col_scale((float)cmax[(size_t)c]);

    // Blocks: maximal runs of RF / precession intervals of the node range.
    std::vector<int64_t> pf, pl;
    std::vector<uint8_t> rf;
    for (int64_t a = r.node_first; a < r.node_last;) {
        const bool isrf = sq.exc[(size_t)a] != 0;
        int64_t b = a + 1;
        while (b < r.node_last && (sq.exc[(size_t)b] != 0) == isrf) ++b;
        pf.push_back(a);
        pl.push_back(b);
        rf.push_back(isrf ? 1 : 0);
        a = b;
    }
    SeqHost sh;
    sh.nn = sq.nn;
    sh.gamma = sq.gamma;
    sh.dt = sq.dt.data();
    sh.gx = sq.gx.data();
    sh.gy = sq.gy.data();
    sh.gz = sq.gz.data();
    sh.coef = sq.coef.data();
    sh.adc = sq.adc.data();
    build_tables(sh, j.tl, j.ps, j.NI, j.mflags, pf.data(), pl.data(), rf.data(), (int64_t)pf.size(),
                 j.segs, j.ts);
    j.rows = 0;
    j.max_rf_adc = 0;
    j.any_fallback = false;
    int64_t max_nsb = 0;
    for (const Segment &g : j.segs) {
        j.rows += g.nadc;
        if (g.rf) j.max_rf_adc = std::max(j.max_rf_adc, g.nadc);
        else {
            max_nsb = std::max(max_nsb, g.nsb);
            j.any_fallback = j.any_fallback || (g.fallback && g.nadc > 0);
        }
    }
    // Chunk: whole anchor blocks, about step_budget spin-samples.
    const int64_t per_sb = j.ns * kBlockTiles * 16;
    int64_t csb = std::max<int64_t>(s->cfg.step_budget / std::max<int64_t>(per_sb, 1), 1);
    csb = std::max<int64_t>(csb / kBlocksPerAnchor, 1) * kBlocksPerAnchor;
    const int64_t cap = (max_nsb + kBlocksPerAnchor - 1) / kBlocksPerAnchor * kBlocksPerAnchor;
    j.chunk_sb = std::max<int64_t>(std::min(csb, cap), kBlocksPerAnchor);

    const int64_t ng = j.ngroups, nmh = std::max(j.ps.nmh, 1);
    const int64_t chunk_tiles = j.chunk_sb * kBlockTiles;
    Layout L;
    j.o_mxy = L.add(16 * j.ns);
    j.o_mz = L.add(8 * j.ns);
    j.o_b64 = L.add(8 * (ng + 1) * 64);
    j.o_aho = L.add(16 * ng * nmh * 64);
    j.o_cdec = L.add(4 * ng * 16);
    j.o_bop = L.add(16 * ng * j.NCB * 2 * 64);
    j.o_csc = L.add(8 * 8 * j.NCB);
    j.o_th0 = L.add(8 * ng * 16);
    j.o_lm = L.add(4 * ng * 16);
    j.o_m0 = L.add(4);
    j.o_seed = L.add(8 * ((j.chunk_sb / kBlocksPerAnchor) * ng * 16 + 16));
    j.o_ba = j.any_fallback ? L.add(8 * ng * j.NI * 64) : -1;
    j.o_kr = j.any_fallback ? L.add(8 * (int64_t)j.ts.kr.size()) : -1;
    j.o_kref = L.add(8 * (int64_t)j.ts.kref.size());
    j.o_bho = L.add(16 * (int64_t)j.ts.bho.size());
    j.o_anc = L.add(8 * (int64_t)j.ts.anc.size());
    j.o_tk = L.add(4 * (int64_t)j.ts.tk.size());
    j.o_kr0 = L.add(8 * (int64_t)j.ts.kr0.size());
    j.o_kend = L.add(8 * (int64_t)j.ts.kend.size());
    j.o_part = L.add(4 * j.nstiles * chunk_tiles * 16 * 16 * j.NCB);
    j.o_sig = L.add(16 * j.rows * j.nc);
    if (j.max_rf_adc > 0) {
        j.o_rfst = L.add(16 * j.ns * j.max_rf_adc);
        j.rf_part_cap = coil_partial_elements(j.ns, j.max_rf_adc, j.nc);
        j.o_rfpart = L.add(16 * j.rf_part_cap);
    }
    j.dev_bytes = L.total;
    j.host_bytes = align_up(16 * j.rows * j.nc);
    if (j.dev_bytes > s->dev_arena.size)
        return fail(s, HOM_ERR_MEMORY, "The request needs more device memory than the server's arena.");
    if (j.host_bytes > s->host_arena.size)
        return fail(s, HOM_ERR_MEMORY, "The request's signal does not fit the server's host arena.");
    return HOM_OK;
}

void fill_view(hom_server *s, Job &j) {
    hom_plan &v = j.view;
    const int64_t f = j.req.spin_first;
    v.device = s->cfg.device;
    v.precision = HOM_F64;
    v.coil_mode = HOM_COIL_GEMM;
    v.ns = j.ns;
    v.nc = j.nc;
    v.nn = j.sq->nn;
    v.max_adc = std::max<int64_t>(j.max_rf_adc, 1);
    v.max_rf_adc = j.max_rf_adc;
    v.gamma = j.sq->gamma;
    v.stream = s->stream;
    v.csm_ld = j.ph->ns;
    v.d.x = (void *)(j.ph->arr(0) + f);
    v.d.y = (void *)(j.ph->arr(1) + f);
    v.d.z = (void *)(j.ph->arr(2) + f);
    v.d.rho = (void *)(j.ph->arr(3) + f);
    v.d.T1 = (void *)(j.ph->arr(4) + f);
    v.d.T2 = (void *)(j.ph->arr(5) + f);
    v.d.dw = (void *)(j.ph->arr(6) + f);
    v.d.csm = (void *)(j.cs->d + f);
    const int64_t nn = j.sq->nn;
    v.d.gx = j.sq->d;
    v.d.gy = j.sq->d + nn;
    v.d.gz = j.sq->d + 2 * nn;
    v.d.df = j.sq->d + 3 * nn;
    v.d.psi = j.sq->d + 4 * nn;
    v.d.b1 = j.sq->d + 5 * nn;
    v.d.dt = j.sq->d + 7 * nn;
    v.d.adc = j.sq->dadc;
    v.d.mxy = j.at<void>(j.o_mxy);
    v.d.mz = j.at<void>(j.o_mz);
    v.d.states = j.at<void>(j.o_rfst);
    v.d.partials = j.at<void>(j.o_rfpart);
    v.d.partial_capacity = j.rf_part_cap;
    v.has_phantom = v.has_csm = v.has_sequence = true;
}

// Uploads tables and launches the preparation kernel of a newly admitted job.
int prepare_job(hom_server *s, Job &j) {
    auto up = [&](int64_t off, const void *src, size_t bytes) -> hipError_t {
        if (off < 0 || bytes == 0) return hipSuccess;
        return hipMemcpyAsync(j.dev + off, src, bytes, hipMemcpyHostToDevice, s->stream);
    };
    hipError_t e = hipSuccess;
    const TableSet &t = j.ts;
    if (e == hipSuccess) e = up(j.o_csc, j.csc.data(), 8 * j.csc.size());
    if (e == hipSuccess && j.o_kr >= 0) e = up(j.o_kr, t.kr.data(), 8 * t.kr.size());
    if (e == hipSuccess) e = up(j.o_kref, t.kref.data(), 8 * t.kref.size());
    if (e == hipSuccess) e = up(j.o_bho, t.bho.data(), 16 * t.bho.size());
    if (e == hipSuccess) e = up(j.o_anc, t.anc.data(), 8 * t.anc.size());
    if (e == hipSuccess) e = up(j.o_tk, t.tk.data(), 4 * t.tk.size());
    if (e == hipSuccess) e = up(j.o_kr0, t.kr0.data(), 8 * t.kr0.size());
    if (e == hipSuccess) e = up(j.o_kend, t.kend.data(), 8 * t.kend.size());
    if (e != hipSuccess) return hip_fail(s, e, "table upload");
    const int64_t f = j.req.spin_first;
    hipLaunchKernelGGL(k_srv_prep, dim3(2048), dim3(256), 0, s->stream, j.ns, j.nc, j.ngroups, j.NCB,
                       j.NI, j.ps, j.tl, j.ph->arr(0) + f, j.ph->arr(1) + f, j.ph->arr(2) + f,
                       j.ph->arr(6) + f, j.ph->arr(5) + f, j.ph->arr(3) + f,
                       (const double2 *)(j.cs->d + f), j.ph->ns, j.at<const double>(j.o_csc),
                       j.at<half8_t>(j.o_bop), j.at<double>(j.o_b64), j.at<half8_t>(j.o_aho),
                       j.at<float>(j.o_cdec), j.at<double>(j.o_ba), j.at<double2>(j.o_mxy),
                       j.at<double>(j.o_mz));
    return hip_fail(s, hipGetLastError(), "request preparation");
}

void add_list(ListDesc &L, int slot, int64_t blocks) {
    L.job[L.n] = slot;
    L.prefix[L.n + 1] = L.prefix[L.n] + blocks;
    ++L.n;
}

struct StepPlan {
    std::vector<MainDesc> mains;  // item_lo holds the entry's item count until finish_mains
    int slots = 0;
};

// Writes the fused-kernel entries with item offsets into the descriptor.
void finish_mains(StepDesc &d, const StepPlan &sp) {
    int64_t acc = 0;
    for (size_t i = 0; i < sp.mains.size(); ++i) {
        d.main[i] = sp.mains[i];
        d.main[i].item_lo = acc;
        acc += sp.mains[i].item_lo;
    }
    d.n_main = (int32_t)sp.mains.size();
    d.total_items = acc;
}

// Fills the descriptor with one chunk of job j; returns false when the job
// has no precession work left this step (RF-only tail or finished).
bool add_chunk(hom_server *s, Job &j, StepDesc &d, StepPlan &sp) {
    if (j.seg >= j.segs.size()) return false;
    const Segment &g = j.segs[j.seg];
    const int slot = sp.slots;
    const int64_t f = j.req.spin_first;
    JobDesc &J = d.jobs[slot];
    std::memset(&J, 0, sizeof(J));
    J.ns = j.ns;
    J.ngroups = j.ngroups;
    J.nc = j.nc;
    J.nstiles = j.nstiles;
    J.x = j.ph->arr(0) + f;
    J.y = j.ph->arr(1) + f;
    J.z = j.ph->arr(2) + f;
    J.rho = j.ph->arr(3) + f;
    J.T1 = j.ph->arr(4) + f;
    J.T2 = j.ph->arr(5) + f;
    J.dw = j.ph->arr(6) + f;
    J.mxy = j.at<double2>(j.o_mxy);
    J.mz = j.at<double>(j.o_mz);
    J.th0 = j.at<double>(j.o_th0);
    J.lm = j.at<float>(j.o_lm);
    J.m0max = j.at<float>(j.o_m0);
    J.csc = j.at<double>(j.o_csc);
    J.seed = j.at<double>(j.o_seed);
    J.ps = j.ps;
    J.tl = j.tl;
    J.kend = j.at<double>(j.o_kend) + g.kend_off;
    J.T = g.T;
    sp.slots++;
    if (g.nadc == 0) {
        add_list(d.end, slot, (j.ns + 255) / 256);
        j.seg++;
        j.sb_next = 0;
        return true;
    }
    const int64_t sb_lo = j.sb_next, sb_hi = std::min(sb_lo + j.chunk_sb, g.nsb);
    const int64_t tile_lo = sb_lo * kBlockTiles, tile_hi = std::min(sb_hi * kBlockTiles, g.tiles);
    const int64_t ab_lo = sb_lo / kBlocksPerAnchor;
    const int64_t ab_hi = (sb_hi + kBlocksPerAnchor - 1) / kBlocksPerAnchor;
    J.do_seg = sb_lo == 0;
    J.do_seed = !g.fallback;
    if (J.do_seg) add_list(d.m0, slot, (j.ns + 255) / 256);
    if (J.do_seed) {
        J.anc = j.at<double>(j.o_anc) + g.anc_off + ab_lo * j.ps.nho;
        J.nab = ab_hi - ab_lo;
    }
    if (J.do_seg || J.do_seed) add_list(d.ss, slot, (j.ngroups * 16 + 255) / 256);
    for (int cb0 = 0; cb0 < j.NCB; cb0 += 4) {
        const int ncb = std::min(4, j.NCB - cb0);
        MainDesc m{};
        MainOps &o = m.o;
        o.ngroups = j.ngroups;
        o.gpt = j.gpt;
        o.cb0 = cb0;
        o.ncb_total = j.NCB;
        o.tile_lo = tile_lo;
        o.tile_hi = tile_hi;
        o.sb_lo = sb_lo;
        o.ab_lo = ab_lo;
        o.basis64 = j.at<double>(j.o_b64);
        o.aho = j.at<half8_t>(j.o_aho);
        o.cdec = j.at<float>(j.o_cdec);
        o.lm = j.at<float>(j.o_lm);
        o.seed = j.at<double>(j.o_seed);
        o.bop = j.at<half8_t>(j.o_bop);
        o.kref = j.at<double>(j.o_kref) + g.kref_off;
        o.bho = j.at<half8_t>(j.o_bho) + g.bho_off;
        o.tk = j.at<float>(j.o_tk) + g.s_off;
        o.partials = j.at<float>(j.o_part);
        o.basisA = j.at<double>(j.o_ba);
        o.th0 = j.at<double>(j.o_th0);
        o.kr = j.o_kr >= 0 ? j.at<double>(j.o_kr) + g.kr_off : nullptr;
        m.nsb = sb_hi - sb_lo;
        m.item_lo = m.nsb * j.nstiles;  // item count; turned into offsets later
        m.cls = g.fallback ? fp64_class(j.NI, ncb) : split_class(j.ps.nmh, ncb);
        sp.mains.push_back(m);
    }
    J.nadc = std::min(tile_hi * 16, g.nadc) - tile_lo * 16;
    J.rows = (tile_hi - tile_lo) * 16;
    J.cols = 16 * (int64_t)j.NCB;
    J.partials = j.at<float>(j.o_part);
    J.kr0 = j.at<double>(j.o_kr0) + g.s_off + tile_lo * 16;
    J.sig = j.at<double2>(j.o_sig) + g.row0 + tile_lo * 16;
    J.ld = std::max<int64_t>(j.rows, 1);
    add_list(d.fin, slot, (J.nadc * j.nc + 255) / 256);
    if (sb_hi == g.nsb) {
        add_list(d.end, slot, (j.ns + 255) / 256);
        j.seg++;
        j.sb_next = 0;
    } else {
        j.sb_next = sb_hi;
    }
    return true;
}

void enqueue_prologue(hom_server *s) {
    const StepDesc *sd = s->d_desc;
    float *m0acc = &s->d_desc->m0acc[0];
    hipLaunchKernelGGL(k_srv_m0max, dim3(kLightGrid), dim3(256), 0, s->stream, sd, m0acc);
    hipLaunchKernelGGL(k_srv_segseed, dim3(kLightGrid), dim3(256), 0, s->stream, sd,
                       (const float *)m0acc);
}

void enqueue_fused(hom_server *s, int grid) {
    hipLaunchKernelGGL(k_srv_fused, dim3((unsigned)grid), dim3(256), 0, s->stream, s->d_desc,
                       &s->d_desc->ctr[0]);
}

// The step's fixed chain of four kernels on the server stream.
void enqueue_kernels(hom_server *s) {
    enqueue_prologue(s);
    enqueue_fused(s, s->cfg.main_grid);
    hipLaunchKernelGGL(k_srv_finend, dim3(kLightGrid), dim3(256), 0, s->stream, s->d_desc);
}

int launch_step(hom_server *s) {
    if (!s->cfg.use_graphs) {
        enqueue_kernels(s);
        return hip_fail(s, hipGetLastError(), "step launch");
    }
    const uint32_t key = (uint32_t)s->cfg.main_grid;
    auto it = s->graphs.find(key);
    if (it == s->graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) return hip_fail(s, e, "hipStreamBeginCapture");
        enqueue_kernels(s);
        e = hipStreamEndCapture(s->stream, &g);
        if (e != hipSuccess) return hip_fail(s, e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(s, e, "hipGraphInstantiate");
        it = s->graphs.emplace(key, ex).first;
        s->graphs_captured++;
    }
    hipError_t e = hipGraphLaunch(it->second, s->stream);
    if (e != hipSuccess) return hip_fail(s, e, "hipGraphLaunch");
    s->graph_launches++;
    return HOM_OK;
}

void release_resident_locked(hom_server *s, Job &j) {
    auto drop = [&](auto &map, auto *obj, auto freer) {
        if (!obj) return;
        obj->refs--;
        if (obj->released && obj->refs == 0) {
            for (auto it = map.begin(); it != map.end(); ++it)
                if (it->second.get() == obj) {
                    freer(obj);
                    map.erase(it);
                    break;
                }
        }
    };
    drop(s->phantoms, j.ph, [](ResPhantom *p) { (void)hipFree(p->d); });
    drop(s->csms, j.cs, [](ResCsm *c) { (void)hipFree(c->d); });
    drop(s->seqs, j.sq, [](ResSeq *q) { (void)hipFree(q->d); (void)hipFree(q->dadc); });
    j.ph = nullptr;
    j.cs = nullptr;
    j.sq = nullptr;
}

// Frees a job's arenas (caller holds mu; the device is idle for it).
void free_job_memory_locked(hom_server *s, Job &j) {
    if (j.dev_off >= 0) s->dev_arena.release(j.dev_off, j.dev_bytes);
    j.dev_off = -1;
    j.dev = nullptr;
    s->stats.device_bytes_used = s->dev_arena.used;
}

// After a device error or a failed step the server stops: every request that
// has not finished fails -- the active ones release their arena memory, the
// queued ones (which hold no arena memory yet) their resident pins and host
// tables -- and every waiter wakes.
void fail_unfinished_locked(hom_server *s, const std::string &msg) {
    const double t = now_seconds();
    for (Job *j : s->active) {
        j->state = HOM_REQ_FAILED;
        j->err = msg;
        j->t_done = t;
        free_job_memory_locked(s, *j);
        if (j->host_off >= 0) s->host_arena.release(j->host_off, j->host_bytes);
        j->host_off = -1;
        j->result = nullptr;
        release_resident_locked(s, *j);
        s->stats.failed++;
    }
    s->active.clear();
    for (Job *j : s->queue) {
        j->state = HOM_REQ_FAILED;
        j->err = msg;
        j->t_done = t;
        release_resident_locked(s, *j);
        j->ts = TableSet{};
        s->stats.failed++;
    }
    s->queue.clear();
    s->stats.active = 0;
    s->stats.queued = 0;
    s->cv_done.notify_all();
}

// One scheduler step.  Runs under step_mu (never under mu while on the GPU).
int step_once(hom_server *s, int *worked) {
    *worked = 0;
    if (const hipError_t e = hipSetDevice(s->cfg.device); e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
    std::vector<Job *> admitted;
    std::vector<Job *> act;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->fatal) return fail_locked(s, HOM_ERR_DEVICE, "The server stopped after a device error.");
        while (!s->queue.empty() && (int)s->active.size() < s->cfg.max_active) {
            Job *j = s->queue.front();
            const int64_t off = s->dev_arena.alloc(j->dev_bytes);
            if (off < 0) break;
            const int64_t hoff = s->host_arena.alloc(j->host_bytes);
            if (hoff < 0) {
                s->dev_arena.release(off, j->dev_bytes);
                break;
            }
            s->queue.pop_front();
            j->dev_off = off;
            j->dev = s->dev_base + off;
            j->host_off = hoff;
            j->result = reinterpret_cast<double2 *>(s->host_base + hoff);
            j->state = HOM_REQ_ACTIVE;
            j->t_admit = now_seconds();
            s->active.push_back(j);
            admitted.push_back(j);
        }
        s->stats.device_bytes_used = s->dev_arena.used;
        s->stats.device_bytes_peak = std::max(s->stats.device_bytes_peak, s->dev_arena.used);
        s->stats.queued = (int64_t)s->queue.size();
        s->stats.active = (int64_t)s->active.size();
        s->stats.max_active_seen = std::max<int64_t>(s->stats.max_active_seen, (int64_t)s->active.size());
        act = s->active;
    }
    if (act.empty()) return HOM_OK;
    *worked = 1;
    const double t0 = now_seconds();
    int rc = HOM_OK;
    for (Job *j : admitted) {
        fill_view(s, *j);
        rc = prepare_job(s, *j);
        if (rc) break;
    }
    // RF blocks due now, then one chunk of every job with precession work.
    StepDesc &d = *s->h_desc;
    std::memset(&d, 0, offsetof(StepDesc, jobs));
    StepPlan sp;
    std::vector<Job *> finishing;
    for (Job *j : act) {
        if (rc) break;
        while (j->seg < j->segs.size() && j->segs[j->seg].rf) {
            const Segment &g = j->segs[j->seg];
            hipError_t e = launch_plain_excitation(&j->view, g.first, g.last, j->flags);
            if (e == hipSuccess && g.nadc > 0)
                e = launch_coil_reduce(&j->view, j->view.d.states, j->ns, g.nadc,
                                       j->at<double2>(j->o_sig) + g.row0, std::max<int64_t>(j->rows, 1));
            if (e != hipSuccess) {
                rc = hip_fail(s, e, "RF block");
                break;
            }
            j->seg++;
        }
        if (rc) break;
        const int need = (j->NCB + 3) / 4;
        if (j->seg < j->segs.size() && (sp.slots >= kMaxJobs || (int)sp.mains.size() + need > kMaxMain))
            continue;  // descriptor full: this job waits for the next step
        if (!add_chunk(s, *j, d, sp)) {
            finishing.push_back(j);
            continue;
        }
        // done after this step when no segment is left (RF tails run next step)
        if (j->seg >= j->segs.size()) finishing.push_back(j);
    }
    if (!rc && (sp.slots > 0)) {
        finish_mains(d, sp);
        const size_t bytes = offsetof(StepDesc, main) + sizeof(MainDesc) * std::max<size_t>(sp.mains.size(), 1);
        hipError_t e = hipMemcpyAsync(s->d_desc, s->h_desc, bytes, hipMemcpyHostToDevice, s->stream);
        rc = e != hipSuccess ? hip_fail(s, e, "descriptor upload") : launch_step(s);
    }
    if (!rc)
        for (Job *j : finishing) {
            hipError_t e = hipSuccess;
            if (j->rows > 0 && j->nc > 0)
                e = hipMemcpyAsync(j->result, j->at<double2>(j->o_sig), (size_t)(16 * j->rows * j->nc),
                                   hipMemcpyDeviceToHost, s->stream);
            if (e != hipSuccess) {
                rc = hip_fail(s, e, "result download");
                break;
            }
        }
    if (!rc) rc = hip_fail(s, hipStreamSynchronize(s->stream), "step");
    const double t1 = now_seconds();
    std::lock_guard<std::mutex> lk(s->mu);
    if (!rc && s->test_fail_step > 0 && s->stats.steps + 1 == s->test_fail_step)
        rc = fail_locked(s, HOM_ERR_DEVICE, "step: injected device error (HOM_SERVE_TEST_FAIL_STEP)");
    s->stats.steps++;
    s->stats.busy_seconds += t1 - t0;
    if (rc) {
        s->fatal = true;
        fail_unfinished_locked(s, s->err);
        return rc;
    }
    for (Job *j : finishing) {
        j->state = HOM_REQ_DONE;
        j->t_done = t1;
        free_job_memory_locked(s, *j);
        release_resident_locked(s, *j);
        j->ts = TableSet{};
        s->stats.completed++;
        s->stats.spin_samples += (double)j->ns * (double)j->rows;
        s->active.erase(std::find(s->active.begin(), s->active.end(), j));
    }
    s->stats.active = (int64_t)s->active.size();
    if (!finishing.empty()) {
        s->epoch++;
        s->cv_done.notify_all();
    }
    return HOM_OK;
}

// step_once behind a catch: a C++ exception in a step (a failed host
// allocation) stops the server like a device error: the active requests fail
// and waiters wake.
int step_guarded(hom_server *s, int *worked) {
    std::string msg;
    int code = HOM_ERR_STATE;
    try {
        return step_once(s, worked);
    } catch (const std::bad_alloc &) {
        msg = "The server stopped: host memory allocation failed in a step.";
        code = HOM_ERR_MEMORY;
    } catch (const std::exception &e) {
        msg = std::string("The server stopped: ") + e.what();
    } catch (...) {
        msg = "The server stopped after an internal error.";
    }
    *worked = 0;
    std::lock_guard<std::mutex> lk(s->mu);
    s->fatal = true;
    s->err = msg;
    fail_unfinished_locked(s, msg);
    s->cv_done.notify_all();
    return code;
}

// Runs f() for a server entry point (s is NULL while a server is created):
// a failed host allocation becomes HOM_ERR_MEMORY, any other exception
// HOM_ERR_STATE, so no C++ exception crosses the C ABI.
template <typename F>
int guarded_server(hom_server *s, F &&f) {
    std::string msg;
    int code = HOM_ERR_STATE;
    try {
        return f();
    } catch (const std::bad_alloc &) {
        msg = "Host memory allocation failed.";
        code = HOM_ERR_MEMORY;
    } catch (const std::exception &e) {
        msg = std::string("Internal error: ") + e.what();
    } catch (...) {
        msg = "Internal error.";
    }
    if (!s) {
        g_server_create_error = msg;
        return code;
    }
    return fail(s, code, msg);
}

void worker_main(hom_server *s) {
    (void)hipSetDevice(s->cfg.device);  // each step sets and checks it again
    uint64_t blocked_at = ~(uint64_t)0;  // epoch at which the queue head could not be admitted
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                return s->stop || !s->active.empty() || (!s->queue.empty() && s->epoch != blocked_at);
            });
            if (s->stop) return;
            if (s->fatal) {
                s->cv_work.wait(lk, [&] { return s->stop; });
                return;
            }
        }
        std::lock_guard<std::mutex> step(s->step_mu);
        uint64_t e0 = 0;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            e0 = s->epoch;
        }
        int worked = 0;
        step_guarded(s, &worked);
        blocked_at = worked ? ~(uint64_t)0 : e0;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->blocked = !worked && !s->queue.empty() && s->epoch == e0;
        }
        if (!worked) s->cv_done.notify_all();
    }
}

int64_t default_grid() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            cus = v;
    }
    return 3 * (int64_t)cus;
}

}  // namespace

extern "C" {

const char *hom_server_last_error(const hom_server *s) {
    if (!s) return g_server_create_error.c_str();
    return s->err.c_str();
}

static int hom_server_create_impl(hom_server **out, const hom_server_config *config) {
    g_server_create_error.clear();
    if (!out) {
        g_server_create_error = "The output pointer is NULL.";
        return HOM_ERR_ARGUMENT;
    }
    *out = nullptr;
    hom_server_config c{};
    if (config) c = *config;
    if (c.device_bytes == 0) c.device_bytes = (int64_t)8 << 30;
    if (c.host_bytes == 0) c.host_bytes = (int64_t)1 << 30;
    if (c.max_active == 0) c.max_active = HOM_SERVE_MAX_ACTIVE;
    if (c.step_budget == 0) c.step_budget = (int64_t)1 << 33;
    if (c.gpt == 0) c.gpt = 256;
    auto bad = [&](const char *m) {
        g_server_create_error = m;
        return HOM_ERR_ARGUMENT;
    };
    if (c.device_bytes < 0 || c.host_bytes < 0) return bad("Arena sizes must be positive.");
    if (c.max_active < 1 || c.max_active > HOM_SERVE_MAX_ACTIVE)
        return bad("max_active must be between 1 and HOM_SERVE_MAX_ACTIVE.");
    if (c.step_budget < 1) return bad("The step budget must be positive.");
    if (c.gpt != 64 && c.gpt != 128 && c.gpt != 256 && c.gpt != 512)
        return bad("gpt must be 64, 128, 256 or 512.");
    if (c.main_grid < 0 || c.main_grid % 8) return bad("main_grid must be a nonnegative multiple of 8.");
    if (c.use_graphs != 0 && c.use_graphs != 1) return bad("use_graphs must be 0 or 1.");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        g_server_create_error = "No HIP device is available.";
        return HOM_ERR_DEVICE;
    }
    if (c.device < 0 || c.device >= count) return bad("The device index is out of range.");
    if (hipSetDevice(c.device) != hipSuccess) {
        g_server_create_error = "hipSetDevice failed.";
        return HOM_ERR_DEVICE;
    }
    if (c.main_grid == 0) c.main_grid = (int)default_grid();
    std::unique_ptr<hom_server> s(new (std::nothrow) hom_server());
    if (!s) {
        g_server_create_error = "Out of host memory.";
        return HOM_ERR_MEMORY;
    }
    s->cfg = c;
    auto dev_fail = [&](const char *m) {
        g_server_create_error = m;
        if (s->stream) (void)hipStreamDestroy(s->stream);
        if (s->dev_base) (void)hipFree(s->dev_base);
        if (s->host_base) (void)hipHostFree(s->host_base);
        if (s->h_desc) (void)hipHostFree(s->h_desc);
        if (s->d_desc) (void)hipFree(s->d_desc);
        return HOM_ERR_MEMORY;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return dev_fail("hipStreamCreate failed.");
    if (hom::dev_malloc((void **)&s->dev_base, (size_t)c.device_bytes) != hipSuccess)
        return dev_fail("Device arena allocation failed.");
    if (hipHostMalloc((void **)&s->host_base, (size_t)c.host_bytes, hipHostMallocDefault) != hipSuccess)
        return dev_fail("Pinned host arena allocation failed.");
    if (hipHostMalloc((void **)&s->h_desc, sizeof(StepDesc), hipHostMallocDefault) != hipSuccess)
        return dev_fail("Pinned descriptor allocation failed.");
    if (hom::dev_malloc((void **)&s->d_desc, sizeof(StepDesc)) != hipSuccess)
        return dev_fail("Device descriptor allocation failed.");
    std::memset(s->h_desc, 0, sizeof(StepDesc));
    s->dev_arena.init(c.device_bytes);
    s->host_arena.init(c.host_bytes);
    if (const char *v = std::getenv("HOM_SERVE_TEST_FAIL_STEP")) s->test_fail_step = std::atoll(v);
    hom_server *raw = s.release();
    if (c.threaded) raw->worker = std::thread(worker_main, raw);
    *out = raw;
    return HOM_OK;
}

int hom_server_destroy(hom_server *s) {
    if (!s) return HOM_ERR_ARGUMENT;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    // releases proceed regardless of earlier device errors
    (void)hipSetDevice(s->cfg.device);
    (void)hipStreamSynchronize(s->stream);
    for (auto &kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto &kv : s->phantoms) (void)hipFree(kv.second->d);
    for (auto &kv : s->csms) (void)hipFree(kv.second->d);
    for (auto &kv : s->seqs) {
        (void)hipFree(kv.second->d);
        (void)hipFree(kv.second->dadc);
    }
    (void)hipStreamDestroy(s->stream);
    (void)hipFree(s->dev_base);
    (void)hipHostFree(s->host_base);
    (void)hipHostFree(s->h_desc);
    (void)hipFree(s->d_desc);
    delete s;
    return HOM_OK;
}

int hom_server_get_config(hom_server *s, hom_server_config *out) {
    if (!s || !out) return HOM_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->cfg;
    return HOM_OK;
}

static int hom_server_put_phantom_impl(hom_server *s, int64_t ns, const double *x, const double *y,
                                       const double *z, const double *rho, const double *T1, const double *T2,
                                       const double *dw, int64_t *id) {
    if (!s) return HOM_ERR_ARGUMENT;
    if (!id) return fail(s, HOM_ERR_ARGUMENT, "The id pointer is NULL.");
    if (ns < 1) return fail(s, HOM_ERR_DIMENSION, "A phantom needs at least one spin.");
    const double *src[7] = {x, y, z, rho, T1, T2, dw};
    for (const double *p : src)
        if (!p) return fail(s, HOM_ERR_ARGUMENT, "Every phantom array needs a host pointer.");
    for (int64_t i = 0; i < ns; ++i)
        if (!(T1[i] > 0.0) || !(T2[i] > 0.0))
            return fail(s, HOM_ERR_ARGUMENT, "T1 and T2 must be positive.");
    if (const hipError_t e = hipSetDevice(s->cfg.device); e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
    auto ph = std::make_unique<ResPhantom>();
    ph->ns = ns;
    if (hom::dev_malloc((void **)&ph->d, sizeof(double) * 7 * (size_t)ns) != hipSuccess)
        return fail(s, HOM_ERR_MEMORY, "Device allocation failed for the phantom.");
    for (int i = 0; i < 7; ++i) {
        hipError_t e = hipMemcpy(ph->d + (size_t)i * ns, src[i], sizeof(double) * ns, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(ph->d);
            return hip_fail(s, e, "phantom upload");
        }
    }
    ph->hx.assign(x, x + ns);
    ph->hy.assign(y, y + ns);
    ph->hz.assign(z, z + ns);
    ph->hdw.assign(dw, dw + ns);
    const ResPhantom &phr = *ph;
    ph->bmax.build(ns, 17, [&](int64_t i, int64_t j) { return basis_abs(phr, i, j); });
    std::lock_guard<std::mutex> lk(s->mu);
    *id = s->next_id++;
    s->phantoms[*id] = std::move(ph);
    return HOM_OK;
}

static int hom_server_put_csm_impl(hom_server *s, int64_t phantom, int64_t nc, const double *csm, int64_t ld,
                                   int64_t *id) {
    if (!s) return HOM_ERR_ARGUMENT;
    if (!id || !csm) return fail(s, HOM_ERR_ARGUMENT, "A CSM pointer is NULL.");
    int64_t ns = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->phantoms.find(phantom);
        if (it == s->phantoms.end() || it->second->released)
            return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown phantom id.");
        ns = it->second->ns;
    }
    if (nc < 1) return fail(s, HOM_ERR_DIMENSION, "A CSM needs at least one coil.");
    if (ld < ns) return fail(s, HOM_ERR_DIMENSION, "The CSM leading dimension is smaller than the spin count.");
    if (const hipError_t e = hipSetDevice(s->cfg.device); e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
    auto cs = std::make_unique<ResCsm>();
    cs->phantom = phantom;
    cs->nc = nc;
    cs->ns = ns;
    cs->host.resize((size_t)(2 * ns * nc));
    for (int64_t c = 0; c < nc; ++c)
        std::memcpy(&cs->host[(size_t)(2 * c * ns)], csm + 2 * c * ld, sizeof(double) * 2 * (size_t)ns);
    if (hom::dev_malloc((void **)&cs->d, sizeof(double2) * (size_t)(ns * nc)) != hipSuccess)
        return fail(s, HOM_ERR_MEMORY, "Device allocation failed for the CSM.");
    hipError_t e = hipMemcpy(cs->d, cs->host.data(), sizeof(double2) * (size_t)(ns * nc), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(cs->d);
        return hip_fail(s, e, "CSM upload");
    }
    const ResCsm &csr = *cs;
    cs->cmax2.build(ns, nc, [&](int64_t i, int64_t c) { return csm_abs2(csr, i, c); });
    std::lock_guard<std::mutex> lk(s->mu);
    *id = s->next_id++;
    s->csms[*id] = std::move(cs);
    return HOM_OK;
}

static int hom_server_put_sequence_impl(hom_server *s, const hom_sequence_host *seq, double gamma, int64_t *id) {
    if (!s) return HOM_ERR_ARGUMENT;
    if (!id || !seq) return fail(s, HOM_ERR_ARGUMENT, "A sequence pointer is NULL.");
    const int64_t n = seq->n_nodes, ni = n - 1;
    if (n < 2) return fail(s, HOM_ERR_DIMENSION, "A sequence needs at least two nodes.");
    if (!seq->t || !seq->gx || !seq->gy || !seq->gz || !seq->b1 || !seq->df || !seq->psi ||
        !seq->adc || !seq->dt || !seq->excitation)
        return fail(s, HOM_ERR_ARGUMENT, "Every sequence array needs a host pointer.");
    for (int j = 0; j < 16; ++j)
        if (!seq->h[j]) return fail(s, HOM_ERR_ARGUMENT, "All sixteen coefficient waveforms are required.");
    for (int64_t i = 0; i < ni; ++i)
        if (!(seq->dt[i] >= 0.0)) return fail(s, HOM_ERR_ARGUMENT, "Time steps must be nonnegative.");
    if (!(gamma > 0.0)) return fail(s, HOM_ERR_ARGUMENT, "gamma must be positive.");
    if (const hipError_t e = hipSetDevice(s->cfg.device); e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
    auto q = std::make_unique<ResSeq>();
    q->nn = n;
    q->gamma = gamma;
    q->t.assign(seq->t, seq->t + n);
    q->dt.assign(seq->dt, seq->dt + ni);
    q->gx.assign(seq->gx, seq->gx + n);
    q->gy.assign(seq->gy, seq->gy + n);
    q->gz.assign(seq->gz, seq->gz + n);
    q->adc.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) q->adc[(size_t)i] = seq->adc[i] ? 1 : 0;
    q->exc.resize((size_t)ni);
    for (int64_t i = 0; i < ni; ++i) q->exc[(size_t)i] = seq->excitation[i] ? 1 : 0;
    q->coef.resize((size_t)(16 * n));
    for (int j = 0; j < 16; ++j) std::copy(seq->h[j], seq->h[j] + n, q->coef.begin() + j * n);
    // device: gx, gy, gz, df, psi, b1 (2n), dt (n - 1, padded to n)
    std::vector<double> h((size_t)(8 * n), 0.0);
    std::copy(seq->gx, seq->gx + n, h.begin());
    std::copy(seq->gy, seq->gy + n, h.begin() + n);
    std::copy(seq->gz, seq->gz + n, h.begin() + 2 * n);
    std::copy(seq->df, seq->df + n, h.begin() + 3 * n);
    std::copy(seq->psi, seq->psi + n, h.begin() + 4 * n);
    std::copy(seq->b1, seq->b1 + 2 * n, h.begin() + 5 * n);
    std::copy(seq->dt, seq->dt + ni, h.begin() + 7 * n);
    if (hom::dev_malloc((void **)&q->d, sizeof(double) * h.size()) != hipSuccess ||
        hom::dev_malloc((void **)&q->dadc, (size_t)n) != hipSuccess) {
        if (q->d) (void)hipFree(q->d);
        return fail(s, HOM_ERR_MEMORY, "Device allocation failed for the sequence.");
    }
    hipError_t e = hipMemcpy(q->d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(q->dadc, q->adc.data(), (size_t)n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(q->d);
        (void)hipFree(q->dadc);
        return hip_fail(s, e, "sequence upload");
    }
    std::lock_guard<std::mutex> lk(s->mu);
    *id = s->next_id++;
    s->seqs[*id] = std::move(q);
    return HOM_OK;
}

static int hom_server_release_impl(hom_server *s, int kind, int64_t id) {
    if (!s) return HOM_ERR_ARGUMENT;
    if (const hipError_t e = hipSetDevice(s->cfg.device); e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
    std::lock_guard<std::mutex> lk(s->mu);
    auto go = [&](auto &map, auto freer) -> int {
        auto it = map.find(id);
        if (it == map.end() || it->second->released) return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown resident id.");
        it->second->released = true;
        if (it->second->refs == 0) {
            freer(it->second.get());
            map.erase(it);
        }
        return HOM_OK;
    };
    switch (kind) {
        case HOM_RES_PHANTOM:
            for (auto &kv : s->csms)
                if (kv.second->phantom == id && !kv.second->released)
                    return fail_locked(s, HOM_ERR_STATE, "Release the phantom's CSMs first.");
            return go(s->phantoms, [](ResPhantom *p) { (void)hipFree(p->d); });
        case HOM_RES_CSM: return go(s->csms, [](ResCsm *c) { (void)hipFree(c->d); });
        case HOM_RES_SEQUENCE:
            return go(s->seqs, [](ResSeq *q) {
                (void)hipFree(q->d);
                (void)hipFree(q->dadc);
            });
        default: return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown resident kind.");
    }
}

static int hom_server_footprint_impl(hom_server *s, const hom_request *req, int64_t *bytes) {
    if (!s) return HOM_ERR_ARGUMENT;
    if (!req || !bytes) return fail(s, HOM_ERR_ARGUMENT, "A pointer is NULL.");
    Job j;
    int gpt = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        const int rc = resolve_job_locked(s, *req, j);
        if (rc) return rc;
        gpt = s->cfg.gpt;
    }
    const int rc = plan_job(s, j, gpt);
    std::lock_guard<std::mutex> lk(s->mu);
    release_resident_locked(s, j);
    if (rc) return rc;
    *bytes = j.dev_bytes;
    return HOM_OK;
}

static int hom_server_submit_impl(hom_server *s, const hom_request *req, int64_t *rid) {
    if (!s) return HOM_ERR_ARGUMENT;
    if (!req || !rid) return fail(s, HOM_ERR_ARGUMENT, "A pointer is NULL.");
    const double t = now_seconds();
    auto j = std::make_unique<Job>();
    int gpt = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->fatal) return fail_locked(s, HOM_ERR_DEVICE, "The server stopped after a device error.");
        const int rc = resolve_job_locked(s, *req, *j);
        if (rc) return rc;
        gpt = s->cfg.gpt;
    }
    // Host tables are built without the lock, so the scheduler never waits on them.
    const int rc = plan_job(s, *j, gpt);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (rc) {
            release_resident_locked(s, *j);
            return rc;
        }
        j->t_submit = t;
        j->id = s->next_id++;
        *rid = j->id;
        s->queue.push_back(j.get());
        s->jobs[j->id] = std::move(j);
        s->stats.submitted++;
        s->stats.queued = (int64_t)s->queue.size();
        s->epoch++;
    }
    s->cv_work.notify_one();
    return HOM_OK;
}

static int hom_server_step_impl(hom_server *s, int *worked) {
    if (!s) return HOM_ERR_ARGUMENT;
    if (s->cfg.threaded) return fail(s, HOM_ERR_STATE, "The worker thread runs the steps of this server.");
    int w = 0;
    std::lock_guard<std::mutex> step(s->step_mu);
    const int rc = step_guarded(s, &w);
    if (worked) *worked = w;
    return rc;
}

static int hom_server_drain_impl(hom_server *s) {
    if (!s) return HOM_ERR_ARGUMENT;
    if (!s->cfg.threaded) {
        for (;;) {
            int w = 0;
            const int rc = hom_server_step(s, &w);
            if (rc) return rc;
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->queue.empty() && s->active.empty()) return HOM_OK;
            if (!w)
                return fail_locked(s, HOM_ERR_MEMORY,
                                   "Queued requests cannot be admitted: fetch finished results to free the host arena.");
        }
    }
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        return s->fatal || s->blocked || (s->queue.empty() && s->active.empty());
    });
    if (s->fatal) return fail_locked(s, HOM_ERR_DEVICE, "The server stopped after a device error.");
    if (!s->queue.empty() && s->active.empty() && s->blocked)
        return fail_locked(s, HOM_ERR_MEMORY,
                           "Queued requests cannot be admitted: fetch finished results to free the host arena.");
    return HOM_OK;
}

int hom_server_status(hom_server *s, int64_t rid, int *state) {
    if (!s || !state) return HOM_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(rid);
    if (it == s->jobs.end()) return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown request id.");
    *state = it->second->state;
    return HOM_OK;
}

int hom_server_wait(hom_server *s, int64_t rid, double timeout_s, int *state) {
    if (!s) return HOM_ERR_ARGUMENT;
    if (!s->cfg.threaded) {
        for (;;) {
            int st = 0;
            int rc = hom_server_status(s, rid, &st);
            if (rc) return rc;
            if (st == HOM_REQ_DONE || st == HOM_REQ_FAILED) {
                if (state) *state = st;
                return HOM_OK;
            }
            int w = 0;
            rc = hom_server_step(s, &w);
            if (rc) return rc;
            if (!w)
                return fail(s, HOM_ERR_MEMORY,
                            "The request cannot be admitted: fetch finished results to free the host arena.");
        }
    }
    std::unique_lock<std::mutex> lk(s->mu);
    // The request is looked up again on every wake-up: another thread may
    // fetch its result (which forgets it) while this one waits.
    auto find = [&]() -> Job * {
        auto it = s->jobs.find(rid);
        return it == s->jobs.end() ? nullptr : it->second.get();
    };
    if (!find()) return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown request id.");
    auto ready = [&] {
        const Job *j = find();
        return !j || j->state == HOM_REQ_DONE || j->state == HOM_REQ_FAILED || s->fatal ||
               (j->state == HOM_REQ_QUEUED && s->blocked && s->active.empty());
    };
    if (timeout_s < 0) s->cv_done.wait(lk, ready);
    else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
    const Job *j = find();
    if (!j) return fail_locked(s, HOM_ERR_ARGUMENT, "The request's result was fetched by another call.");
    if (state) *state = j->state;
    if (j->state == HOM_REQ_QUEUED && s->blocked && s->active.empty())
        return fail_locked(s, HOM_ERR_MEMORY,
                           "The request cannot be admitted: fetch finished results to free the host arena.");
    return HOM_OK;
}

int hom_server_result_shape(hom_server *s, int64_t rid, int64_t *rows, int64_t *cols) {
    if (!s || !rows || !cols) return HOM_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(rid);
    if (it == s->jobs.end()) return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown request id.");
    *rows = it->second->rows;
    *cols = it->second->nc;
    return HOM_OK;
}

static int hom_server_result_impl(hom_server *s, int64_t rid, double *out, int64_t ld) {
    if (!s) return HOM_ERR_ARGUMENT;
    const double2 *src = nullptr;
    int64_t rows = 0, nc = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(rid);
        if (it == s->jobs.end() || it->second->fetching)
            return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown request id.");
        Job &j = *it->second;
        if (j.state == HOM_REQ_FAILED) {
            const std::string m = "The request failed: " + j.err;
            if (j.host_off >= 0) s->host_arena.release(j.host_off, j.host_bytes);
            s->jobs.erase(it);
            return fail_locked(s, HOM_ERR_DEVICE, m);
        }
        if (j.state != HOM_REQ_DONE) return fail_locked(s, HOM_ERR_STATE, "The request has not finished.");
        if (j.rows > 0 && (!out || ld < j.rows))
            return fail_locked(s, HOM_ERR_DIMENSION,
                               "The output is NULL or its leading dimension is too small.");
        j.fetching = true;
        src = j.result;
        rows = j.rows;
        nc = j.nc;
    }
    // The copy runs outside the lock so the scheduler is never held up by it.
    for (int64_t c = 0; c < nc; ++c)
        std::memcpy(out + 2 * c * ld, src + c * rows, sizeof(double2) * (size_t)rows);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(rid);
    s->host_arena.release(it->second->host_off, it->second->host_bytes);
    s->jobs.erase(it);
    s->epoch++;
    s->cv_work.notify_one();
    return HOM_OK;
}

int hom_server_times(hom_server *s, int64_t rid, double *t_submit, double *t_admit, double *t_done) {
    if (!s) return HOM_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(rid);
    if (it == s->jobs.end()) return fail_locked(s, HOM_ERR_ARGUMENT, "Unknown request id.");
    if (t_submit) *t_submit = it->second->t_submit;
    if (t_admit) *t_admit = it->second->t_admit;
    if (t_done) *t_done = it->second->t_done;
    return HOM_OK;
}

static int hom_server_get_stats_impl(hom_server *s, hom_server_stats *out) {
    if (!s || !out) return HOM_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->graph_launches = s->graph_launches.load();
    out->graphs_captured = s->graphs_captured.load();
    out->device_bytes_used = s->dev_arena.used;
    out->device_bytes_peak = s->dev_arena.peak;
    return HOM_OK;
}

static int hom_server_autotune_impl(hom_server *s, const hom_request *probe, const int *gpt_list, int n_gpt,
                                    const int *grid_list, int n_grid, int *best_gpt, int *best_grid,
                                    double *best_seconds) {
    if (!s) return HOM_ERR_ARGUMENT;
    if (!probe) return fail(s, HOM_ERR_ARGUMENT, "The probe request is NULL.");
    static const int kGpt[] = {64, 128, 256, 512};
    const int cus = (int)(default_grid() / 3);
    std::vector<int> gpts(gpt_list ? gpt_list : kGpt, gpt_list ? gpt_list + n_gpt : kGpt + 4);
    std::vector<int> grids;
    if (grid_list) grids.assign(grid_list, grid_list + n_grid);
    else grids = {2 * cus, 3 * cus, 4 * cus, 6 * cus};
    for (int g : gpts)
        if (g != 64 && g != 128 && g != 256 && g != 512) return fail(s, HOM_ERR_ARGUMENT, "gpt must be 64, 128, 256 or 512.");
    for (int g : grids)
        if (g <= 0 || g % 8) return fail(s, HOM_ERR_ARGUMENT, "Grid sizes must be positive multiples of 8.");
    if (gpts.empty() || grids.empty()) return fail(s, HOM_ERR_ARGUMENT, "The candidate lists are empty.");
    std::lock_guard<std::mutex> step(s->step_mu);
    if (const hipError_t e = hipSetDevice(s->cfg.device); e != hipSuccess) return hip_fail(s, e, "hipSetDevice");
    Job j;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->queue.empty() || !s->active.empty())
            return fail_locked(s, HOM_ERR_STATE, "The server must be idle to autotune.");
        const int rc = resolve_job_locked(s, *probe, j);
        if (rc) return rc;
    }
    {
        // smallest gpt: most spin tiles, so the partial buffer fits every candidate
        const int rc = plan_job(s, j, *std::min_element(gpts.begin(), gpts.end()));
        std::lock_guard<std::mutex> lk(s->mu);
        if (!rc) j.dev_off = s->dev_arena.alloc(j.dev_bytes);
        if (rc || j.dev_off < 0) {
            release_resident_locked(s, j);
            return rc ? rc : fail_locked(s, HOM_ERR_MEMORY, "The probe does not fit the device arena.");
        }
        j.dev = s->dev_base + j.dev_off;
    }
    int rc = HOM_OK;
    double best = -1.0;
    int bg = gpts[0], bgrid = grids[0];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    fill_view(s, j);
    rc = prepare_job(s, j);
    // first precession segment with samples
    while (!rc && j.seg < j.segs.size() && (j.segs[j.seg].rf || j.segs[j.seg].nadc == 0)) {
        const Segment &g = j.segs[j.seg];
        if (g.rf) {
            hipError_t e = launch_plain_excitation(&j.view, g.first, g.last, j.flags);
            if (e == hipSuccess && g.nadc > 0)
                e = launch_coil_reduce(&j.view, j.view.d.states, j.ns, g.nadc, j.at<double2>(j.o_sig) + g.row0,
                                       std::max<int64_t>(j.rows, 1));
            if (e != hipSuccess) rc = hip_fail(s, e, "RF block");
        }
        j.seg++;
    }
    if (!rc && j.seg >= j.segs.size()) rc = fail(s, HOM_ERR_ARGUMENT, "The probe has no precession samples.");
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(s, HOM_ERR_DEVICE, "hipEventCreate failed.");
    const size_t seg0 = j.seg;
    for (int g : gpts) {
        if (rc) break;
        for (int grid : grids) {
            if (rc) break;
            j.gpt = g;
            j.nstiles = std::max<int64_t>((j.ngroups + g - 1) / g, 1);
            j.seg = seg0;
            j.sb_next = 0;
            StepDesc &d = *s->h_desc;
            std::memset(&d, 0, offsetof(StepDesc, jobs));
            StepPlan sp;
            add_chunk(s, j, d, sp);
            finish_mains(d, sp);
            double t_min = -1.0;
            for (int rep = 0; rep < 3 && !rc; ++rep) {
                std::memset(&d.ctr, 0, sizeof(d.ctr));
                hipError_t e = hipMemcpyAsync(s->d_desc, s->h_desc, sizeof(StepDesc), hipMemcpyHostToDevice, s->stream);
                if (e == hipSuccess) {
                    // prep kernels of the chunk (untimed), then the fused kernel (timed)
                    enqueue_prologue(s);
                    e = hipEventRecord(e0, s->stream);
                    if (e == hipSuccess) {
                        enqueue_fused(s, grid);
                        e = hipEventRecord(e1, s->stream);
                    }
                    const hipError_t es = hipStreamSynchronize(s->stream);
                    if (e == hipSuccess) e = es;
                }
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e != hipSuccess) {
                    rc = hip_fail(s, e, "autotune run");
                    break;
                }
                if (t_min < 0 || ms < t_min) t_min = ms;
            }
            if (!rc && (best < 0 || t_min < best)) {
                best = t_min;
                bg = g;
                bgrid = grid;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    const hipError_t es = hipStreamSynchronize(s->stream);
    std::lock_guard<std::mutex> lk(s->mu);
    s->dev_arena.release(j.dev_off, j.dev_bytes);
    release_resident_locked(s, j);
    if (rc) return rc;
    if (es != hipSuccess)
        return fail_locked(s, HOM_ERR_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(es));
    s->cfg.gpt = bg;
    s->cfg.main_grid = bgrid;
    for (auto &kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
    s->graphs.clear();
    if (best_gpt) *best_gpt = bg;
    if (best_grid) *best_grid = bgrid;
    if (best_seconds) *best_seconds = best * 1e-3;
    return HOM_OK;
}

// Entry points that allocate host memory or start threads: no exception
// crosses the C ABI.
int hom_server_put_phantom(hom_server *s, int64_t ns, const double *x, const double *y,
                           const double *z, const double *rho, const double *T1, const double *T2,
                           const double *dw, int64_t *id) {
    return guarded_server(s, [&] {
        return hom_server_put_phantom_impl(s, ns, x, y, z, rho, T1, T2, dw, id);
    });
}

int hom_server_put_csm(hom_server *s, int64_t phantom, int64_t nc, const double *csm, int64_t ld,
                       int64_t *id) {
    return guarded_server(s, [&] {
        return hom_server_put_csm_impl(s, phantom, nc, csm, ld, id);
    });
}

int hom_server_put_sequence(hom_server *s, const hom_sequence_host *seq, double gamma, int64_t *id) {
    return guarded_server(s, [&] {
        return hom_server_put_sequence_impl(s, seq, gamma, id);
    });
}

int hom_server_release(hom_server *s, int kind, int64_t id) {
    return guarded_server(s, [&] {
        return hom_server_release_impl(s, kind, id);
    });
}

int hom_server_footprint(hom_server *s, const hom_request *req, int64_t *bytes) {
    return guarded_server(s, [&] {
        return hom_server_footprint_impl(s, req, bytes);
    });
}

int hom_server_submit(hom_server *s, const hom_request *req, int64_t *rid) {
    return guarded_server(s, [&] {
        return hom_server_submit_impl(s, req, rid);
    });
}

int hom_server_step(hom_server *s, int *worked) {
    return guarded_server(s, [&] {
        return hom_server_step_impl(s, worked);
    });
}

int hom_server_drain(hom_server *s) {
    return guarded_server(s, [&] {
        return hom_server_drain_impl(s);
    });
}

int hom_server_result(hom_server *s, int64_t rid, double *out, int64_t ld) {
    return guarded_server(s, [&] {
        return hom_server_result_impl(s, rid, out, ld);
    });
}

int hom_server_get_stats(hom_server *s, hom_server_stats *out) {
    return guarded_server(s, [&] {
        return hom_server_get_stats_impl(s, out);
    });
}

int hom_server_autotune(hom_server *s, const hom_request *probe, const int *gpt_list, int n_gpt,
                        const int *grid_list, int n_grid, int *best_gpt, int *best_grid,
                        double *best_seconds) {
    return guarded_server(s, [&] {
        return hom_server_autotune_impl(s, probe, gpt_list, n_gpt, grid_list, n_grid, best_gpt, best_grid, best_seconds);
    });
}

int hom_server_create(hom_server **out, const hom_server_config *config) {
    return guarded_server(nullptr, [&] {
        return hom_server_create_impl(out, config);
    });
}

}  // extern "C"
