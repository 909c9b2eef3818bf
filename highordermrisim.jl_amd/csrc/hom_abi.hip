// hom_abi.hip -- C ABI of libhom.so (declared in include/hom_abi.h).
//
// Plan lifecycle, host->device residency, the per-block drop-in entry points
// and the whole-run block loop.  Replaces, on the device side, the reference's
// HighOrderGPUWorkspace / run_high_order_{precession,excitation}! seam
// (src/simulation/kernels/HighOrderGPU.jl:16-58,124-140,223-233) and the loop
// of run_sim_time_iter! (src/simulation/SimulatorCore.jl:96-152).
#include "hom_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>

using namespace hom;

namespace {
thread_local std::string g_create_error;

template <typename T>
int upload_vec(hom_plan *p, void **dst, const double *src, int64_t n) {
    if (n == 0) return HOM_OK;
    if (!*dst) {
        hipError_t e = hom::dev_malloc(dst, sizeof(T) * (size_t)n);
        if (e != hipSuccess) return set_error(p, HOM_ERR_MEMORY, "hipMalloc failed for an input array");
    }
    const size_t bytes = sizeof(T) * (size_t)n;
    if (std::is_same<T, double>::value) return staged_upload(p, *dst, bytes, src, bytes, bytes, 1);
    std::vector<T> tmp((size_t)n);
    for (int64_t i = 0; i < n; ++i) tmp[(size_t)i] = (T)src[i];
    return staged_upload(p, *dst, bytes, tmp.data(), bytes, bytes, 1);
}

int upload_real(hom_plan *p, void **dst, const double *src, int64_t n) {
    return p->is_f32() ? upload_vec<float>(p, dst, src, n) : upload_vec<double>(p, dst, src, n);
}

int ensure_alloc(hom_plan *p, void **ptr, size_t bytes) {
    if (*ptr || bytes == 0) return HOM_OK;
    if (hom::dev_malloc(ptr, bytes) != hipSuccess) {
        *ptr = nullptr;
        return set_error(p, HOM_ERR_MEMORY, "device allocation failed");
    }
    return HOM_OK;
}

// Device words for the upload statistics: max(17, Nc).
int ensure_stats(hom_plan *p) {
    return ensure_alloc(p, (void **)&p->d.stats,
                        sizeof(unsigned long long) * (size_t)std::max<int64_t>(17, p->nc));
}

int ensure_partial_elements(hom_plan *p, int64_t need) {
    if (need <= p->d.partial_capacity) return HOM_OK;
    if (p->d.partials) (void)hipFree(p->d.partials);
    p->d.partials = nullptr;
    p->d.partial_capacity = 0;
    if (hom::dev_malloc(&p->d.partials, (size_t)need * p->complex_size()) != hipSuccess) {
        p->d.partials = nullptr;
        return set_error(p, HOM_ERR_MEMORY, "device allocation failed for coil partials");
    }
    p->d.partial_capacity = need;
    return HOM_OK;
}

int ensure_partials(hom_plan *p, int64_t n_adc) {
    return ensure_partial_elements(p, coil_partial_elements(p->ns, n_adc, p->nc));
}

int require_inputs(hom_plan *p) {
    if (!p->has_phantom) return set_error(p, HOM_ERR_STATE, "The phantom has not been uploaded.");
    if (!p->has_csm) return set_error(p, HOM_ERR_STATE, "The CSM has not been uploaded.");
    if (!p->has_sequence) return set_error(p, HOM_ERR_STATE, "The sequence has not been uploaded.");
    return HOM_OK;
}

int64_t host_block_adc(const hom_plan *p, int64_t first, int64_t last) {
    int64_t n = 0;
    for (int64_t i = first + 1; i <= last; ++i) n += p->h_adc[(size_t)i] ? 1 : 0;
    return n;
}

int check_block(hom_plan *p, int64_t first, int64_t last, bool rf, int64_t *nadc) {
    if (first < 0 || last < first || last >= p->nn)
        return set_error(p, HOM_ERR_ARGUMENT, "The block node range is outside the sequence.");
    for (int64_t i = first; i < last; ++i)
        if ((p->h_exc[(size_t)i] != 0) != rf)
            return set_error(p, HOM_ERR_ARGUMENT,
                             rf ? "An excitation block contains a free-precession interval."
                                : "A precession block contains an RF interval.");
    *nadc = host_block_adc(p, first, last);
    if (*nadc > p->max_adc)
        return set_error(p, HOM_ERR_DIMENSION, "The block acquires more samples than max_adc.");
    // a fused plan's state buffer holds max_rf_adc samples (ADC during RF only)
    if (rf && p->fused() && *nadc > p->max_rf_adc)
        return set_error(p, HOM_ERR_DIMENSION, "The excitation block acquires more samples than max_rf_adc.");
    return HOM_OK;
}

int block_common(hom_plan *p, int64_t first, int64_t last, int flags, void *out, int64_t ld,
                 bool rf) {
    if (!p) return HOM_ERR_ARGUMENT;
    int rc = require_inputs(p);
    if (rc) return rc;
    int64_t nadc = 0;
    rc = check_block(p, first, last, rf, &nadc);
    if (rc) return rc;
    if (nadc > 0 && !out)
        return set_error(p, HOM_ERR_ARGUMENT, "A block with ADC samples needs an output buffer.");
    if (nadc > 0 && ld < nadc)
        return set_error(p, HOM_ERR_DIMENSION, "The output leading dimension is smaller than the block's ADC count.");
    if (!rf && p->precision == HOM_MIXED)
        return mixed_precession_block(p, first, last, flags, out, ld);
    if (!rf && p->fused()) {
        if (nadc > 0 && (rc = ensure_partial_elements(p, fused_partial_elements(p->ns, nadc, p->nc))))
            return rc;
        return check_hip(p, launch_fused_precession(p, first, last, flags, nadc, out, ld),
                         "fused precession kernel");
    }
    if (nadc > 0) {
        rc = ensure_partials(p, nadc);
        rc = rc ? rc : ensure_state_cols(p, nadc);
        if (rc) return rc;
    }
    rc = check_hip(p, rf ? launch_plain_excitation(p, first, last, flags)
                         : launch_plain_precession(p, first, last, flags),
                   rf ? "excitation kernel" : "precession kernel");
    if (rc) return rc;
    if (nadc > 0)
        rc = check_hip(p, launch_coil_reduce(p, p->d.states, p->ns, nadc, out, ld), "coil reduction");
    return rc;
}

void free_ptr(void *&ptr) {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
}
}  // namespace

namespace hom {

int ensure_state_cols(hom_plan *p, int64_t n_adc) {
    if (n_adc <= p->d.state_cols || p->ns == 0) return HOM_OK;
    if (const int e = check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize")) return e;
    free_ptr(p->d.states);
    p->d.state_cols = 0;
    if (dev_malloc(&p->d.states, p->complex_size() * (size_t)p->ns * (size_t)n_adc) != hipSuccess)
        return set_error(p, HOM_ERR_MEMORY, "device allocation failed for ADC states");
    p->d.state_cols = n_adc;
    return HOM_OK;
}

int set_error(hom_plan *p, int code, const std::string &msg) {
    if (p)
        p->err = msg;
    else
        g_create_error = msg;
    return code;
}

int check_hip(hom_plan *p, hipError_t e, const char *what) {
    if (e == hipSuccess) return HOM_OK;
    return set_error(p, HOM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

int enter_device(hom_plan *p, bool sync) {
    int rc = check_hip(p, hipSetDevice(p->device), "hipSetDevice");
    if (!rc && sync && p->stream) rc = check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    return rc;
}

ScopedTimer::ScopedTimer(hom_plan *plan, int kernel_id, hipStream_t stream) : p(plan), id(kernel_id) {
    if (!p || !p->timing) return;
    st = stream ? stream : p->stream;
    TimerSlot &s = p->timers[id];
    if (s.used == s.starts.size()) {
        hipEvent_t a = nullptr, b = nullptr;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
        s.starts.push_back(a);
        s.stops.push_back(b);
    }
    (void)hipEventRecord(s.starts[s.used], st);  // timing only
    stop_ev = s.stops[s.used];
}

ScopedTimer::~ScopedTimer() {
    if (!stop_ev) return;
    (void)hipEventRecord(stop_ev, st);  // timing only
    p->timers[id].used++;
}

}  // namespace hom

namespace {
// Runs f() for a plan entry point (p may be NULL while a plan is created): a
// failed host allocation becomes HOM_ERR_MEMORY, any other exception
// HOM_ERR_STATE with its text in hom_last_error.
template <typename F>
int guarded(hom_plan *p, F &&f) {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return set_error(p, HOM_ERR_MEMORY, "Host memory allocation failed.");
    } catch (const std::exception &e) {
        return set_error(p, HOM_ERR_STATE, std::string("Internal error: ") + e.what());
    } catch (...) {
        return set_error(p, HOM_ERR_STATE, "Internal error.");
    }
}
}  // namespace

extern "C" {

int hom_abi_version(void) { return HOM_ABI_VERSION; }

int hom_device_count(int *count) {
    if (!count) return HOM_ERR_ARGUMENT;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return HOM_OK;
}

const char *hom_last_error(const hom_plan *plan) {
    return plan ? plan->err.c_str() : g_create_error.c_str();
}

static int hom_plan_create_impl(hom_plan **out, int device, int precision, int64_t n_spins, int64_t n_coils,
                                int64_t n_nodes, int64_t max_adc, int64_t max_rf_adc, double gamma,
                                int coil_combination) {
    g_create_error.clear();
    if (!out) return set_error(nullptr, HOM_ERR_ARGUMENT, "The plan output pointer is NULL.");
    *out = nullptr;
    if (max_adc < 0)
        return set_error(nullptr, HOM_ERR_ARGUMENT, "The maximum ADC count must be a nonnegative integer.");
    if (max_rf_adc < 0 || max_rf_adc > max_adc)
        return set_error(nullptr, HOM_ERR_ARGUMENT,
                         "The RF ADC count must be an integer between zero and max_adc.");
    if (coil_combination < HOM_COIL_GEMM || coil_combination > HOM_COIL_FUSED_TILED)
        return set_error(nullptr, HOM_ERR_ARGUMENT, "Unknown coil_combination.");
    if (precision != HOM_F32 && precision != HOM_F64 && precision != HOM_MIXED)
        return set_error(nullptr, HOM_ERR_ARGUMENT, "The precision must be f32, f64 or mixed.");
    if (n_spins < 0) return set_error(nullptr, HOM_ERR_ARGUMENT, "The spin count must be nonnegative.");
    if (n_coils < 1) return set_error(nullptr, HOM_ERR_ARGUMENT, "The CSM must contain at least one coil.");
    if (n_nodes < 1) return set_error(nullptr, HOM_ERR_ARGUMENT, "The sequence must contain at least one node.");
    if (!(gamma > 0.0) || !std::isfinite(gamma))
        return set_error(nullptr, HOM_ERR_ARGUMENT, "gamma must be positive and finite.");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(nullptr, HOM_ERR_DEVICE, "No HIP device is available.");
    if (device < 0 || device >= ndev)
        return set_error(nullptr, HOM_ERR_ARGUMENT, "The device index is out of range.");
    if (hipSetDevice(device) != hipSuccess)
        return set_error(nullptr, HOM_ERR_DEVICE, "hipSetDevice failed.");
    hom_plan *p = new (std::nothrow) hom_plan();
    if (!p) return set_error(nullptr, HOM_ERR_MEMORY, "host allocation failed");
    p->device = device;
    p->precision = precision;
    p->coil_mode = coil_combination;
    p->ns = n_spins;
    p->nc = n_coils;
    p->nn = n_nodes;
    p->max_adc = max_adc;
    p->max_rf_adc = max_rf_adc;
    p->gamma = gamma;
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        delete p;
        return set_error(nullptr, HOM_ERR_DEVICE, "hipStreamCreate failed.");
    }
    if (hipDeviceGetAttribute(&p->cu_count, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        p->cu_count <= 0) {
        (void)hipGetLastError();
        p->cu_count = 256;  // MI355X
    }
    // this device's address of the allocation-free reduction's scratch array
    if (scratch_setup(device) != hipSuccess) {
        (void)hipStreamDestroy(p->stream);
        delete p;
        return set_error(nullptr, HOM_ERR_DEVICE, "hipGetSymbolAddress failed for the reduction scratch array.");
    }
    const size_t R = p->real_size(), Cb = p->complex_size();
    const size_t ns = (size_t)n_spins;
    int rc = HOM_OK;
    rc |= ensure_alloc(p, &p->d.mxy, Cb * ns);
    rc |= ensure_alloc(p, &p->d.mz, R * ns);
    // ADC states: every block's in the per-block path (allocated here, as the
    // reference's workspace does); only those acquired during RF when the
    // fused kernel (f32/f64 fused modes) or the segment kernels (mixed) take
    // the precession blocks.  A mixed plan grows the buffer on demand when a
    // precession block is stepped through the per-block entry point.
    const int64_t cols = (p->fused() || precision == HOM_MIXED) ? max_rf_adc : max_adc;
    rc |= ensure_alloc(p, &p->d.states, Cb * ns * (size_t)cols);
    if (!rc) p->d.state_cols = cols;
    if (rc) {
        std::string msg = p->err;
        hom_plan_destroy(p);
        return set_error(nullptr, HOM_ERR_MEMORY, msg.empty() ? "device allocation failed" : msg);
    }
    *out = p;
    return HOM_OK;
}

int hom_check_workspace_options(int device, int precision, int coil_combination, int64_t groupsize,
                                int64_t adc_tile, int64_t coil_tile) {
    g_create_error.clear();
    if (coil_combination < HOM_COIL_GEMM || coil_combination > HOM_COIL_FUSED_TILED)
        return set_error(nullptr, HOM_ERR_ARGUMENT, "Unknown coil_combination.");
    if (precision != HOM_F32 && precision != HOM_F64 && precision != HOM_MIXED)
        return set_error(nullptr, HOM_ERR_ARGUMENT, "The precision must be f32, f64 or mixed.");
    const bool fused = coil_combination == HOM_COIL_FUSED_WARP || coil_combination == HOM_COIL_FUSED_TILED;
    if (!fused) return HOM_OK;
    if (!(groupsize >= 32 && groupsize <= 1024 && groupsize % 32 == 0))
        return set_error(nullptr, HOM_ERR_ARGUMENT, "CUDA fused groupsize must be a multiple of 32, at most 1024.");
    const int64_t A = coil_combination == HOM_COIL_FUSED_WARP ? 1 : adc_tile;
    const int64_t C = coil_combination == HOM_COIL_FUSED_WARP ? 1 : coil_tile;
    if (!((A == 1 || A == 2 || A == 4 || A == 8) && (C == 1 || C == 2 || C == 4)))
        return set_error(nullptr, HOM_ERR_ARGUMENT,
                         "Fused ADC/coil tiles must be selected from 1/2/4/8 and 1/2/4.");
    int64_t shared = 64 * 1024;
    if (device >= 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(nullptr, HOM_ERR_DEVICE, "hipDeviceGetAttribute failed.");
        }
        shared = v;
    }
    const int64_t csize = precision == HOM_F32 ? 8 : 16;
    if (groupsize * A * csize > shared)
        return set_error(nullptr, HOM_ERR_ARGUMENT, "The fused ADC tile exceeds the device shared-memory limit.");
    return HOM_OK;
}

int hom_plan_destroy(hom_plan *p) {
    if (!p) return HOM_OK;
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);  // releases proceed regardless
    mixed_free(p);
    staging_release(p);
    void **all[] = {&p->d.x,  &p->d.y,  &p->d.z,  &p->d.rho, &p->d.T1, &p->d.T2,  &p->d.dw,
                    &p->d.csm, &p->d.coef, &p->d.gx, &p->d.gy, &p->d.gz, &p->d.dt, &p->d.b1,
                    &p->d.df, &p->d.psi, &p->d.mxy, &p->d.mz, &p->d.states, &p->d.partials,
                    (void **)&p->d.stats, &p->signal};
    for (void **q : all) free_ptr(*q);
    if (p->d.adc) (void)hipFree(p->d.adc);
    for (void *b : p->user_buffers) (void)hipFree(b);
    for (auto &s : p->timers) {
        for (auto e : s.starts) (void)hipEventDestroy(e);
        for (auto e : s.stops) (void)hipEventDestroy(e);
    }
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return HOM_OK;
}

static int hom_upload_phantom_impl(hom_plan *p, const double *x, const double *y, const double *z,
                                   const double *rho, const double *T1, const double *T2, const double *dw) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (p->ns > 0 && (!x || !y || !z || !rho || !T1 || !T2 || !dw))
        return set_error(p, HOM_ERR_ARGUMENT, "Every phantom field needs a host pointer.");
    if (const int e = enter_device(p, true)) return e;
    int rc = HOM_OK;
    rc = rc ? rc : upload_real(p, &p->d.x, x, p->ns);
    rc = rc ? rc : upload_real(p, &p->d.y, y, p->ns);
    rc = rc ? rc : upload_real(p, &p->d.z, z, p->ns);
    rc = rc ? rc : upload_real(p, &p->d.rho, rho, p->ns);
    rc = rc ? rc : upload_real(p, &p->d.T1, T1, p->ns);
    rc = rc ? rc : upload_real(p, &p->d.T2, T2, p->ns);
    rc = rc ? rc : upload_real(p, &p->d.dw, dw, p->ns);
    if (rc) return rc;
    // max |basis term j| over the spins, reduced on the device
    unsigned long long bits[17] = {};
    if (p->ns > 0) {
        rc = ensure_stats(p);
        rc = rc ? rc : check_hip(p, hipMemsetAsync(p->d.stats, 0, sizeof(bits), p->stream), "hipMemsetAsync");
        rc = rc ? rc : check_hip(p, launch_basis_absmax(p, p->d.stats), "basis statistics");
        rc = rc ? rc : check_hip(p, hipMemcpyAsync(bits, p->d.stats, sizeof(bits), hipMemcpyDeviceToHost,
                                                   p->stream), "hipMemcpyAsync");
        rc = rc ? rc : check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize");
        if (rc) return rc;
    }
    p->h_basis_max[0] = 0.0;
    for (int j = 1; j <= 16; ++j) std::memcpy(&p->h_basis_max[j], &bits[j], sizeof(double));
    p->has_phantom = true;
    if (p->mixed) mixed_free(p);  // cached operands depend on the phantom
    return hom_reset_state(p);
}

static int hom_upload_csm_impl(hom_plan *p, const double *csm, int64_t ld) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (ld < std::max<int64_t>(p->ns, 1))
        return set_error(p, HOM_ERR_DIMENSION, "The CSM leading dimension is smaller than the spin count.");
    if (p->ns > 0 && !csm) return set_error(p, HOM_ERR_ARGUMENT, "The CSM host pointer is NULL.");
    if (const int e = enter_device(p, true)) return e;
    const size_t n = (size_t)(p->ns * p->nc);
    p->h_csm_colmax.assign((size_t)p->nc, 0.0);
    int rc = ensure_alloc(p, &p->d.csm, p->complex_size() * n);
    if (rc) return rc;
    if (n) {
        if (p->is_f32()) {
            std::vector<float> tmp(2 * n);
            for (int64_t c = 0; c < p->nc; ++c)
                for (int64_t s = 0; s < p->ns; ++s) {
                    const double *v = csm + 2 * (s + c * ld);
                    tmp[2 * (s + c * p->ns)] = (float)v[0];
                    tmp[2 * (s + c * p->ns) + 1] = (float)v[1];
                }
            rc = staged_upload(p, p->d.csm, 8 * n, tmp.data(), 8 * n, 8 * n, 1);
        } else {
            // straight from the caller's columns (column stride ld) into stride ns
            rc = staged_upload(p, p->d.csm, 16 * (size_t)p->ns, csm, 16 * (size_t)ld, 16 * (size_t)p->ns,
                               (size_t)p->nc);
        }
        // column maxima of |csm| for the mixed path's power-of-two scales
        std::vector<unsigned long long> bits((size_t)p->nc, 0ull);
        rc = rc ? rc : ensure_stats(p);
        rc = rc ? rc : check_hip(p, hipMemsetAsync(p->d.stats, 0, sizeof(unsigned long long) * bits.size(),
                                                   p->stream), "hipMemsetAsync");
        rc = rc ? rc : check_hip(p, launch_csm_colmax2(p, p->d.stats), "CSM statistics");
        rc = rc ? rc : check_hip(p, hipMemcpyAsync(bits.data(), p->d.stats, sizeof(unsigned long long) * bits.size(),
                                                   hipMemcpyDeviceToHost, p->stream), "hipMemcpyAsync");
        rc = rc ? rc : check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize");
        if (rc) return rc;
        for (int64_t c = 0; c < p->nc; ++c) {
            double m2;
            std::memcpy(&m2, &bits[(size_t)c], sizeof(double));
            p->h_csm_colmax[(size_t)c] = std::sqrt(m2);
        }
    }
    if (rc) return rc;
    p->has_csm = true;
    if (p->mixed) mixed_free(p);  // cached operands depend on the CSM
    return HOM_OK;
}

static int hom_upload_sequence_impl(hom_plan *p, const hom_sequence_host *seq) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (!seq) return set_error(p, HOM_ERR_ARGUMENT, "The sequence pointer is NULL.");
    if (seq->n_nodes != p->nn)
        return set_error(p, HOM_ERR_DIMENSION, "The sequence node count does not match the plan.");
    const int64_t n = p->nn, ni = n - 1;
    if (!seq->t || !seq->gx || !seq->gy || !seq->gz || !seq->b1 || !seq->df || !seq->psi ||
        !seq->adc || (ni > 0 && (!seq->dt || !seq->excitation)))
        return set_error(p, HOM_ERR_ARGUMENT, "Every sequence array needs a host pointer.");
    for (int j = 0; j < 16; ++j)
        if (!seq->h[j]) return set_error(p, HOM_ERR_ARGUMENT, "All sixteen coefficient waveforms are required.");
    for (int64_t i = 0; i < ni; ++i)
        if (!(seq->dt[i] >= 0.0)) return set_error(p, HOM_ERR_ARGUMENT, "Time steps must be nonnegative.");
    if (const int e = enter_device(p, true)) return e;
    p->h_t.assign(seq->t, seq->t + n);
    p->h_dt.assign(seq->dt ? seq->dt : seq->t, (seq->dt ? seq->dt : seq->t) + ni);
    p->h_gx.assign(seq->gx, seq->gx + n);
    p->h_gy.assign(seq->gy, seq->gy + n);
    p->h_gz.assign(seq->gz, seq->gz + n);
    p->h_adc.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) p->h_adc[(size_t)i] = seq->adc[i] ? 1 : 0;
    p->h_exc.resize((size_t)ni);
    for (int64_t i = 0; i < ni; ++i) p->h_exc[(size_t)i] = seq->excitation[i] ? 1 : 0;
    p->h_coef.resize((size_t)(16 * n));
    for (int j = 0; j < 16; ++j) std::copy(seq->h[j], seq->h[j] + n, p->h_coef.begin() + j * n);
    int rc = HOM_OK;
    rc = rc ? rc : upload_real(p, &p->d.gx, seq->gx, n);
    rc = rc ? rc : upload_real(p, &p->d.gy, seq->gy, n);
    rc = rc ? rc : upload_real(p, &p->d.gz, seq->gz, n);
    rc = rc ? rc : upload_real(p, &p->d.df, seq->df, n);
    rc = rc ? rc : upload_real(p, &p->d.psi, seq->psi, n);
    rc = rc ? rc : upload_real(p, &p->d.b1, seq->b1, 2 * n);
    rc = rc ? rc : upload_real(p, &p->d.coef, p->h_coef.data(), 16 * n);
    if (ni > 0) rc = rc ? rc : upload_real(p, &p->d.dt, seq->dt, ni);
    if (rc) return rc;
    if (!p->d.adc && hom::dev_malloc((void **)&p->d.adc, (size_t)n) != hipSuccess) {
        p->d.adc = nullptr;
        return set_error(p, HOM_ERR_MEMORY, "device allocation failed");
    }
    rc = check_hip(p, hipMemcpy(p->d.adc, p->h_adc.data(), (size_t)n, hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    p->has_sequence = true;
    if (p->mixed) mixed_free(p);  // cached tables depend on the sequence
    return HOM_OK;
}

int hom_reset_state(hom_plan *p) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (!p->has_phantom) return set_error(p, HOM_ERR_STATE, "The phantom has not been uploaded.");
    if (const int e = enter_device(p, false)) return e;
    return check_hip(p, launch_reset_state(p), "state reset");
}

static int hom_set_state_impl(hom_plan *p, const double *mxy, const double *mz) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (p->ns > 0 && (!mxy || !mz)) return set_error(p, HOM_ERR_ARGUMENT, "State pointers are NULL.");
    if (const int e = enter_device(p, true)) return e;
    int rc = upload_real(p, &p->d.mxy, mxy, 2 * p->ns);
    return rc ? rc : upload_real(p, &p->d.mz, mz, p->ns);
}

static int hom_get_state_impl(hom_plan *p, double *mxy, double *mz) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (p->ns > 0 && (!mxy || !mz)) return set_error(p, HOM_ERR_ARGUMENT, "State pointers are NULL.");
    int rc = enter_device(p, true);
    if (rc || p->ns == 0) return rc;
    const size_t ns = (size_t)p->ns;
    if (p->is_f32()) {
        std::vector<float> a(2 * ns), b(ns);
        rc = staged_download(p, a.data(), 8 * ns, p->d.mxy, 8 * ns, 8 * ns, 1);
        if (!rc) rc = staged_download(p, b.data(), 4 * ns, p->d.mz, 4 * ns, 4 * ns, 1);
        if (rc) return rc;
        for (size_t i = 0; i < 2 * ns; ++i) mxy[i] = a[i];
        for (size_t i = 0; i < ns; ++i) mz[i] = b[i];
        return HOM_OK;
    }
    rc = staged_download(p, mxy, 16 * ns, p->d.mxy, 16 * ns, 16 * ns, 1);
    return rc ? rc : staged_download(p, mz, 8 * ns, p->d.mz, 8 * ns, 8 * ns, 1);
}

static int hom_precession_block_impl(hom_plan *p, int64_t first, int64_t last, int flags, void *out, int64_t ld) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (const int e = enter_device(p, false)) return e;
    return block_common(p, first, last, flags, out, ld, false);
}

static int hom_excitation_block_impl(hom_plan *p, int64_t first, int64_t last, int flags, void *out, int64_t ld) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (const int e = enter_device(p, false)) return e;
    return block_common(p, first, last, flags, out, ld, true);
}

static int hom_reduce_signal_impl(hom_plan *p, const void *states, int64_t ld_states, int64_t n_adc, void *out,
                                  int64_t out_ld) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (!p->has_csm) return set_error(p, HOM_ERR_STATE, "The CSM has not been uploaded.");
    if (n_adc < 0) return set_error(p, HOM_ERR_ARGUMENT, "The ADC count must be nonnegative.");
    if (n_adc == 0) return HOM_OK;
    if (!out) return set_error(p, HOM_ERR_ARGUMENT, "The output pointer is NULL.");
    if (out_ld < n_adc) return set_error(p, HOM_ERR_DIMENSION, "The output leading dimension is too small.");
    if (p->ns > 0 && (!states || ld_states < p->ns))
        return set_error(p, HOM_ERR_DIMENSION, "States, CSM, and signal dimensions do not match.");
    if (const int e = enter_device(p, false)) return e;
    int rc = ensure_partials(p, n_adc);
    if (rc) return rc;
    return check_hip(p, launch_coil_reduce(p, states, ld_states, n_adc, out, out_ld), "coil reduction");
}

static int hom_reduce_signal_noalloc_impl(hom_plan *p, const void *states, int64_t ld_states, int64_t n_adc,
                                          void *out, int64_t out_ld) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (!p->has_csm) return set_error(p, HOM_ERR_STATE, "The CSM has not been uploaded.");
    if (n_adc < 0) return set_error(p, HOM_ERR_ARGUMENT, "The ADC count must be nonnegative.");
    if (n_adc == 0) return HOM_OK;
    if (!out) return set_error(p, HOM_ERR_ARGUMENT, "The output pointer is NULL.");
    if (out_ld < n_adc) return set_error(p, HOM_ERR_DIMENSION, "The output leading dimension is too small.");
    if (p->ns > 0 && (!states || ld_states < p->ns))
        return set_error(p, HOM_ERR_DIMENSION, "States, CSM, and signal dimensions do not match.");
    const int64_t elems = coil_partial_elements(p->ns, n_adc, p->nc);
    if (elems > (int64_t)(HOM_REDUCE_SCRATCH_BYTES / p->complex_size()))
        return set_error(p, HOM_ERR_DIMENSION,
                         "The reduction needs more partials than the scratch array holds "
                         "(ceil(Ns / 2048) * n_adc * Nc complex values within HOM_REDUCE_SCRATCH_BYTES).");
    if (const int e = enter_device(p, false)) return e;
    return check_hip(p, launch_coil_reduce_scratch(p, states, ld_states, n_adc, out, out_ld),
                     "coil reduction (scratch)");
}

static int hom_run_impl(hom_plan *p, const int64_t *pf, const int64_t *pl, const uint8_t *rf, int64_t np,
                        int flags) {
    if (!p) return HOM_ERR_ARGUMENT;
    int rc = require_inputs(p);
    if (rc) return rc;
    if (np < 1 || !pf || !pl || !rf)
        return set_error(p, HOM_ERR_ARGUMENT, "The sequence must contain a simulation interval.");
    int64_t total = 0;
    for (int64_t b = 0; b < np; ++b) {
        int64_t nadc = 0;
        rc = check_block(p, pf[b], pl[b], rf[b] != 0, &nadc);
        if (rc) return rc;
        if (b > 0 && pf[b] != pl[b - 1])
            return set_error(p, HOM_ERR_ARGUMENT, "Adjacent parts must share their boundary node.");
        total += nadc;
    }
    if ((rc = enter_device(p, false))) return rc;
    if (total > p->signal_capacity_rows) {
        if ((rc = check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize"))) return rc;
        free_ptr(p->signal);
        p->signal_capacity_rows = 0;
        rc = ensure_alloc(p, &p->signal, p->complex_size() * (size_t)(total * p->nc));
        if (rc) return rc;
        p->signal_capacity_rows = total;
    }
    p->signal_rows = total;
    if (p->precision == HOM_MIXED) return mixed_run(p, pf, pl, rf, np, flags);
    int64_t row = 0;
    for (int64_t b = 0; b < np; ++b) {
        const int64_t nadc = host_block_adc(p, pf[b], pl[b]);
        void *out = nadc ? (char *)p->signal + p->complex_size() * (size_t)row : nullptr;
        rc = block_common(p, pf[b], pl[b], flags, out, std::max<int64_t>(total, 1), rf[b] != 0);
        if (rc) return rc;
        row += nadc;
    }
    return HOM_OK;
}

int hom_synchronize(hom_plan *p) {
    if (!p) return HOM_ERR_ARGUMENT;
    return enter_device(p, true);
}

int hom_signal_rows(hom_plan *p, int64_t *rows) {
    if (!p || !rows) return HOM_ERR_ARGUMENT;
    *rows = p->signal_rows;
    return HOM_OK;
}

int hom_signal_element_size(hom_plan *p, int64_t *bytes) {
    if (!p || !bytes) return HOM_ERR_ARGUMENT;
    *bytes = (int64_t)p->complex_size();
    return HOM_OK;
}

int hom_download_signal(hom_plan *p, void *host, int64_t ld) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (p->signal_rows == 0) return HOM_OK;
    if (!host) return set_error(p, HOM_ERR_ARGUMENT, "The host output pointer is NULL.");
    if (ld < p->signal_rows) return set_error(p, HOM_ERR_DIMENSION, "The host leading dimension is too small.");
    int rc = enter_device(p, true);
    if (rc) return rc;
    const size_t es = p->complex_size();
    return staged_download(p, host, es * (size_t)ld, p->signal, es * (size_t)p->signal_rows,
                           es * (size_t)p->signal_rows, (size_t)p->nc);
}

static int hom_buffer_alloc_impl(hom_plan *p, int64_t bytes, void **dev) {
    if (!p || !dev || bytes < 0) return HOM_ERR_ARGUMENT;
    *dev = nullptr;
    if (const int e = enter_device(p, false)) return e;
    if (bytes == 0) return HOM_OK;
    if (hom::dev_malloc(dev, (size_t)bytes) != hipSuccess) {
        *dev = nullptr;
        return set_error(p, HOM_ERR_MEMORY, "device allocation failed");
    }
    p->user_buffers.push_back(*dev);
    return HOM_OK;
}

int hom_buffer_free(hom_plan *p, void *dev) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (!dev) return HOM_OK;
    auto it = std::find(p->user_buffers.begin(), p->user_buffers.end(), dev);
    if (it == p->user_buffers.end())
        return set_error(p, HOM_ERR_ARGUMENT, "The buffer was not allocated by this plan.");
    const int rc = enter_device(p, true);  // the buffer is released either way
    (void)hipFree(dev);
    p->user_buffers.erase(it);
    return rc;
}

int hom_buffer_upload(hom_plan *p, void *dev, const void *host, int64_t bytes) {
    if (!p || bytes < 0 || (bytes && (!dev || !host))) return HOM_ERR_ARGUMENT;
    if (!bytes) return HOM_OK;
    if (const int e = enter_device(p, true)) return e;
    return staged_upload(p, dev, (size_t)bytes, host, (size_t)bytes, (size_t)bytes, 1);
}

int hom_buffer_download(hom_plan *p, void *host, const void *dev, int64_t bytes) {
    if (!p || bytes < 0 || (bytes && (!dev || !host))) return HOM_ERR_ARGUMENT;
    if (!bytes) return HOM_OK;
    int rc = enter_device(p, true);
    return rc ? rc : staged_download(p, host, (size_t)bytes, dev, (size_t)bytes, (size_t)bytes, 1);
}

static int hom_plan_set_timing_impl(hom_plan *p, int enable) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (const int e = enter_device(p, true)) return e;
    p->timing = enable != 0;
    for (auto &s : p->timers) {
        s.used = 0;
        s.folded_ms = 0.0;
        s.folded_launches = 0;
    }
    return HOM_OK;
}

static int set_spin_tile(hom_plan *p, int groups, bool fixed) {
    if (!p) return HOM_ERR_ARGUMENT;
    if (groups < 16 || groups > 4096 || (groups & (groups - 1)) != 0)
        return set_error(p, HOM_ERR_ARGUMENT, "The spin tile must be a power of two between 16 and 4096 groups.");
    p->spin_tile = groups;
    p->spin_tile_fixed = fixed;
    return HOM_OK;
}

int hom_plan_set_spin_tile(hom_plan *p, int groups) { return set_spin_tile(p, groups, true); }

int hom_plan_set_spin_tile_auto(hom_plan *p, int groups) { return set_spin_tile(p, groups, false); }

int hom_plan_get_spin_tile(const hom_plan *p, int *groups) {
    if (!p || !groups) return HOM_ERR_ARGUMENT;
    *groups = p->spin_tile;
    return HOM_OK;
}

int hom_plan_last_spin_tile(const hom_plan *p, int *groups) {
    if (!p || !groups) return HOM_ERR_ARGUMENT;
    *groups = p->last_spin_tile;
    return HOM_OK;
}

int hom_plan_mixed_stats(const hom_plan *p, int64_t *segments, int64_t *fp64_segments) {
    if (!p || !segments || !fp64_segments) return HOM_ERR_ARGUMENT;
    *segments = p->mixed_segments;
    *fp64_segments = p->mixed_fp64_segments;
    return HOM_OK;
}

static int hom_plan_kernel_time_impl(hom_plan *p, int id, double *total_ms, int64_t *launches) {
    if (!p || id < 0 || id >= HOM_K_COUNT || !total_ms || !launches) return HOM_ERR_ARGUMENT;
    int rc = enter_device(p, true);
    if (rc) return rc;
    TimerSlot &s = p->timers[id];
    for (size_t i = 0; i < s.used; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s.starts[i], s.stops[i]) == hipSuccess) {
            s.folded_ms += ms;
            s.folded_launches += 1;
        }
    }
    s.used = 0;
    *total_ms = s.folded_ms;
    *launches = s.folded_launches;
    return HOM_OK;
}

// Entry points that allocate host memory or start threads: no exception
// crosses the C ABI.
int hom_upload_phantom(hom_plan *p, const double *x, const double *y, const double *z,
                       const double *rho, const double *T1, const double *T2, const double *dw) {
    return guarded(p, [&] {
        return hom_upload_phantom_impl(p, x, y, z, rho, T1, T2, dw);
    });
}

int hom_upload_csm(hom_plan *p, const double *csm, int64_t ld) {
    return guarded(p, [&] {
        return hom_upload_csm_impl(p, csm, ld);
    });
}

int hom_upload_sequence(hom_plan *p, const hom_sequence_host *seq) {
    return guarded(p, [&] {
        return hom_upload_sequence_impl(p, seq);
    });
}

int hom_set_state(hom_plan *p, const double *mxy, const double *mz) {
    return guarded(p, [&] {
        return hom_set_state_impl(p, mxy, mz);
    });
}

int hom_get_state(hom_plan *p, double *mxy, double *mz) {
    return guarded(p, [&] {
        return hom_get_state_impl(p, mxy, mz);
    });
}

int hom_precession_block(hom_plan *p, int64_t first, int64_t last, int flags, void *out, int64_t ld) {
    return guarded(p, [&] {
        return hom_precession_block_impl(p, first, last, flags, out, ld);
    });
}

int hom_excitation_block(hom_plan *p, int64_t first, int64_t last, int flags, void *out, int64_t ld) {
    return guarded(p, [&] {
        return hom_excitation_block_impl(p, first, last, flags, out, ld);
    });
}

int hom_reduce_signal(hom_plan *p, const void *states, int64_t ld_states, int64_t n_adc, void *out,
                      int64_t out_ld) {
    return guarded(p, [&] {
        return hom_reduce_signal_impl(p, states, ld_states, n_adc, out, out_ld);
    });
}

int hom_reduce_signal_noalloc(hom_plan *p, const void *states, int64_t ld_states, int64_t n_adc, void *out,
                              int64_t out_ld) {
    return guarded(p, [&] {
        return hom_reduce_signal_noalloc_impl(p, states, ld_states, n_adc, out, out_ld);
    });
}

int hom_run(hom_plan *p, const int64_t *pf, const int64_t *pl, const uint8_t *rf, int64_t np,
            int flags) {
    return guarded(p, [&] {
        return hom_run_impl(p, pf, pl, rf, np, flags);
    });
}

int hom_buffer_alloc(hom_plan *p, int64_t bytes, void **dev) {
    return guarded(p, [&] {
        return hom_buffer_alloc_impl(p, bytes, dev);
    });
}

int hom_plan_set_timing(hom_plan *p, int enable) {
    return guarded(p, [&] {
        return hom_plan_set_timing_impl(p, enable);
    });
}

int hom_plan_kernel_time(hom_plan *p, int id, double *total_ms, int64_t *launches) {
    return guarded(p, [&] {
        return hom_plan_kernel_time_impl(p, id, total_ms, launches);
    });
}

int hom_plan_create(hom_plan **out, int device, int precision, int64_t n_spins, int64_t n_coils,
                    int64_t n_nodes, int64_t max_adc, int64_t max_rf_adc, double gamma,
                    int coil_combination) {
    return guarded(nullptr, [&] {
        return hom_plan_create_impl(out, device, precision, n_spins, n_coils, n_nodes, max_adc, max_rf_adc, gamma, coil_combination);
    });
}

}  // extern "C"
