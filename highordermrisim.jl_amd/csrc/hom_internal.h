// hom_internal.h -- plan object and kernel launcher interfaces shared by the
// C ABI (hom_abi.cpp) and the HIP kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/hom_abi.h"

namespace hom {

// hipMalloc that leaves nothing behind when it fails: the error is taken out
// of the thread's last-error slot, so a later hipGetLastError() after a
// kernel launch does not report this allocation failure as a launch failure.
template <typename P>
inline hipError_t dev_malloc(P **ptr, size_t bytes) {
    const hipError_t e = hipMalloc(reinterpret_cast<void **>(ptr), bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        (void)hipGetLastError();
    }
    return e;
}

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr int kTermDW = 16;  // term code of the off-resonance term

// Unnormalised real solid harmonic j (1..15) at (x, y, z), in the reference's
// order (src/simulation/kernels/HighOrderGPU.jl:63-71); code kTermDW returns dw.
__host__ __device__ inline double basis_term(int code, double x, double y, double z, double dw) {
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

// Device arrays of one precision (float for HOM_F32, double otherwise).
struct DeviceArrays {
    void *x = nullptr, *y = nullptr, *z = nullptr, *rho = nullptr, *T1 = nullptr,
         *T2 = nullptr, *dw = nullptr;
    void *csm = nullptr;      // Ns x Nc complex, column stride Ns
    void *coef = nullptr;     // 16 x n_nodes, row j contiguous
    void *gx = nullptr, *gy = nullptr, *gz = nullptr, *dt = nullptr, *b1 = nullptr,
         *df = nullptr, *psi = nullptr;
    uint8_t *adc = nullptr;
    void *mxy = nullptr, *mz = nullptr;   // state: Ns complex, Ns real
    void *states = nullptr;               // Ns x state_cols complex (column stride Ns)
    int64_t state_cols = 0;               // ADC samples per spin the state buffer holds
    void *partials = nullptr;             // coil reduction partials
    int64_t partial_capacity = 0;         // in complex elements
    unsigned long long *stats = nullptr;  // upload statistics (max(17, Nc) words)
};

struct MixedWorkspace;  // mixed_kernels.hip

struct TimerSlot {
    std::vector<hipEvent_t> starts, stops;
    size_t used = 0;
    double folded_ms = 0.0;
    int64_t folded_launches = 0;
};

// Two pinned chunk buffers and their copy-done events for staged transfers
// (allocated on the first large copy, released with the plan).
struct Staging {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
};

}  // namespace hom

struct hom_plan {
    int device = 0;
    int precision = HOM_F64;
    int coil_mode = HOM_COIL_GEMM;
    int64_t ns = 0, nc = 0, nn = 0, max_adc = 0, max_rf_adc = 0;
    int64_t csm_ld = 0;  // column stride of d.csm (0: ns)
    double gamma = 0.0;
    hipStream_t stream = nullptr;
    std::string err;
    hom::DeviceArrays d;
    bool has_phantom = false, has_csm = false, has_sequence = false;

    // Host copies used by the mixed path to build its integrated tables.
    std::vector<double> h_t, h_dt, h_gx, h_gy, h_gz, h_coef;  // coef: 16 x nn
    std::vector<uint8_t> h_adc, h_exc;
    std::vector<double> h_csm_colmax;  // max |csm| per coil
    double h_basis_max[17] = {};       // max over spins of |basis term j| (j = 1..16)

    void *signal = nullptr;
    int64_t signal_rows = 0, signal_capacity_rows = 0;

    hom::MixedWorkspace *mixed = nullptr;
    // 16-spin groups per spin tile of the mixed kernels: fixed, or an upper
    // bound of the per-segment choice (hom_plan_set_spin_tile[_auto])
    int spin_tile = 256;
    bool spin_tile_fixed = false;
    int last_spin_tile = 0;  // the longest segment's tile in the last mixed run
    int cu_count = 0;        // compute units of the plan's device
    // Last mixed run: precession segments with samples, and how many of them
    // ran the all-FP64 phase body.
    int64_t mixed_segments = 0, mixed_fp64_segments = 0;

    bool timing = false;
    hom::TimerSlot timers[HOM_K_COUNT];

    std::vector<void *> user_buffers;
    hom::Staging staging;

    bool is_f32() const { return precision == HOM_F32; }
    // f32/f64 plan whose precession blocks run the fused kernel
    bool fused() const {
        return precision != HOM_MIXED && (coil_mode == HOM_COIL_FUSED_WARP || coil_mode == HOM_COIL_FUSED_TILED);
    }
    size_t real_size() const { return is_f32() ? 4 : 8; }
    size_t complex_size() const { return 2 * real_size(); }
};

namespace hom {

// Records a start/stop event pair around one launch when timing is on, on
// the given stream (default: the plan's).
struct ScopedTimer {
    hom_plan *p;
    int id;
    hipStream_t st = nullptr;
    hipEvent_t stop_ev = nullptr;
    ScopedTimer(hom_plan *plan, int kernel_id, hipStream_t stream = nullptr);
    ~ScopedTimer();
};

int set_error(hom_plan *p, int code, const std::string &msg);

// Runs f() for an entry point without a handle: a failed host allocation
// becomes HOM_ERR_MEMORY and any other exception HOM_ERR_STATE, so no C++
// exception crosses the C ABI.
template <typename F>
int guarded_host(F &&f) {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return HOM_ERR_MEMORY;
    } catch (...) {
        return HOM_ERR_STATE;
    }
}
int check_hip(hom_plan *p, hipError_t e, const char *what);
// Grows the ADC-state buffer to hold n_adc samples per spin (after the
// plan's stream has finished with the old one).
int ensure_state_cols(hom_plan *p, int64_t n_adc);
// Makes the plan's device current and, with sync, waits for its stream (so
// that an earlier asynchronous failure is reported here); returns a status.
int enter_device(hom_plan *p, bool sync);

// ---- staged host<->device copies (staging.cpp) ----
// Copies rows x width bytes between pageable host memory (row stride
// host_pitch) and device memory (row stride dev_pitch) on the plan's stream
// and returns when the copy has finished.  Copies of 4 MiB and more go
// through the plan's pinned chunk buffers, packed or unpacked by host
// threads while the previous chunk is in flight; smaller ones (or all of
// them, when pinned memory cannot be had) go through hipMemcpy2DAsync.
int staged_upload(hom_plan *p, void *dev, size_t dev_pitch, const void *host, size_t host_pitch, size_t width,
                  size_t rows);
int staged_download(hom_plan *p, void *host, size_t host_pitch, const void *dev, size_t dev_pitch, size_t width,
                    size_t rows);
void staging_release(hom_plan *p);

// ---- plain kernels (plain_kernels.hip) ----
// Block [first, last] of free precession / excitation for all spins, writing
// ADC states into d.states (column stride Ns).  Returns hipError_t.
hipError_t launch_plain_precession(hom_plan *p, int64_t first, int64_t last, int flags);
hipError_t launch_plain_excitation(hom_plan *p, int64_t first, int64_t last, int flags);
// sig[k, c] = sum_s states[s, k] * csm[s, c] into out (column stride out_ld).
hipError_t launch_coil_reduce(hom_plan *p, const void *states, int64_t ld_states, int64_t n_adc,
                              void *out, int64_t out_ld);
// The same reduction with the stage-1 partials in the library's per-device
// scratch array (HOM_REDUCE_SCRATCH_BYTES) and no kernel timers: no device
// allocation, no HIP event creation.  scratch_setup (device current, once per
// device) looks up that device's address of the array; the launcher returns
// hipErrorNotInitialized when it has not run for the plan's device.
hipError_t scratch_setup(int device);
hipError_t launch_coil_reduce_scratch(hom_plan *p, const void *states, int64_t ld_states,
                                      int64_t n_adc, void *out, int64_t out_ld);
int64_t coil_partial_elements(int64_t ns, int64_t n_adc, int64_t nc);
int64_t fused_partial_elements(int64_t ns, int64_t n_adc, int64_t nc);
hipError_t launch_fused_precession(hom_plan *p, int64_t first, int64_t last, int flags,
                                   int64_t n_adc, void *out, int64_t out_ld);
hipError_t launch_fill_complex(hom_plan *p, void *out, int64_t rows, int64_t cols, int64_t ld);
hipError_t launch_reset_state(hom_plan *p);
// Upload statistics as bit patterns of nonnegative doubles (out zeroed by the
// caller): out[j] = max |basis term j| (j = 1..16); out[c] = max |csm[., c]|^2.
hipError_t launch_basis_absmax(hom_plan *p, unsigned long long *out);
hipError_t launch_csm_colmax2(hom_plan *p, unsigned long long *out);

// ---- mixed path (mixed_kernels.hip) ----
int mixed_run(hom_plan *p, const int64_t *pf, const int64_t *pl, const uint8_t *rf, int64_t np,
              int flags);
// One free-precession block [first, last] of a mixed plan as its own segment
// (states stay on chip; no ADC-state buffer), its Nadc x Nc signal written to
// out (complex double, column stride ld).  The sequence tables are cached for
// the last part list, so stepping block by block rebuilds them per block.
int mixed_precession_block(hom_plan *p, int64_t first, int64_t last, int flags, void *out, int64_t ld);
void mixed_free(hom_plan *p);

}  // namespace hom
