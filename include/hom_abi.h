/*
 * hom_abi.h -- C ABI of the MI355X BlochHighOrder precession + reception
 * library (libhom.so).  Plain pointers and sizes only; every call returns a
 * status (HOM_OK = 0) and the text of the last error is available from
 * hom_last_error().
 *
 * Conventions (Julia layouts, so a ccall passes its arrays unchanged):
 *   - real per-spin and per-node arrays are double on the host;
 *   - complex arrays are interleaved {re, im} of the signal element type;
 *   - matrices are column-major with an explicit leading dimension;
 *   - node indices are 0-based and a block is the inclusive node range
 *     [first, last]; the ADC flag of the shared first node belongs to the
 *     previous block (reference src/simulation/SimulatorCore.jl:115-117).
 *
 * One plan is bound to one device and one HIP stream (a mixed plan runs its
 * main kernels on a second stream of its own, which every call joins back
 * into the plan's stream before it returns) and owns every device
 * allocation it makes.  Calls on one plan must be serialized by the caller;
 * different plans may be used from different threads.  Shapes are validated
 * before any allocation or launch.  No C++ exception leaves the library: a
 * failed host allocation returns HOM_ERR_MEMORY, any other internal failure
 * HOM_ERR_STATE, each with its text in hom_last_error().
 */
#ifndef HOM_ABI_H
#define HOM_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HOM_ABI_VERSION 1

/* Status codes.  The Julia shim maps them to exceptions:
 * ARGUMENT -> ArgumentError, DIMENSION -> DimensionMismatch, rest -> ErrorException. */
enum hom_status {
    HOM_OK = 0,
    HOM_ERR_ARGUMENT = 1,
    HOM_ERR_DIMENSION = 2,
    HOM_ERR_DEVICE = 3,
    HOM_ERR_MEMORY = 4,
    HOM_ERR_STATE = 5
};

/* "precision" of the reference's sim_params (Simulate.jl:90-91, 149-152), plus the
 * mixed-precision fast path (FP64 phase + split-FP16 MFMA coil reduction). */
enum hom_precision { HOM_F32 = 0, HOM_F64 = 1, HOM_MIXED = 2 };

/* "coil_combination" (reference HighOrderGPU.jl:4). */
enum hom_coil_mode {
    HOM_COIL_GEMM = 0,
    HOM_COIL_REDUCTION = 1,
    HOM_COIL_FUSED_WARP = 2,
    HOM_COIL_FUSED_TILED = 3
};

/* BlochHighOrder flags (reference src/simulation/BlochHighOrder.jl:21-43). */
enum hom_flags {
    HOM_HO0 = 1 << 0,
    HOM_HO1 = 1 << 1,
    HOM_HO2 = 1 << 2,
    HOM_HO3 = 1 << 3,
    HOM_DW_EXCITATION = 1 << 4,
    HOM_DW_PRECESSION = 1 << 5
};

/* Kernel ids for hom_plan_kernel_time. */
enum hom_kernel_id {
    HOM_K_PLAIN_PRECESSION = 0,
    HOM_K_PLAIN_EXCITATION = 1,
    HOM_K_COIL_PARTIALS = 2,
    HOM_K_COIL_FINALIZE = 3,
    HOM_K_MIXED_SEGMENT = 4,
    HOM_K_MIXED_FINALIZE = 5,
    HOM_K_MIXED_PREP = 6,
    HOM_K_FUSED_PRECESSION = 7,
    HOM_K_COUNT = 8
};

typedef struct hom_plan hom_plan;

/* Sampled sequence on the host: KomaMRI DiscreteSequence node arrays plus the
 * 16 coefficient waveforms (reference src/datatypes/HO_DiscreteSequence.jl:31-49).
 * t, gx, gy, gz, df, psi, adc and h[j] have n_nodes entries; b1 has n_nodes
 * complex entries; dt and excitation have n_nodes - 1. */
typedef struct hom_sequence_host {
    int64_t n_nodes;
    const double *t, *dt, *gx, *gy, *gz, *b1, *df, *psi;
    const uint8_t *adc, *excitation;
    const double *h[16];
} hom_sequence_host;

int hom_abi_version(void);
int hom_device_count(int *count);

/* Last error of this plan; with plan == NULL, the calling thread's last
 * error from hom_plan_create. */
const char *hom_last_error(const hom_plan *plan);

/* Replaces the HighOrderGPUWorkspace constructor (reference
 * src/simulation/kernels/HighOrderGPU.jl:16-58): same argument checks
 * (0 <= max_rf_adc <= max_adc, known coil mode) plus the device, precision,
 * problem sizes and gamma (Hz/T), which the Julia side carries in its types.
 * As in the reference, an f32/f64 plan in a fused coil mode sizes its
 * ADC-state buffer for max_rf_adc samples: its precession blocks run the
 * fused kernel (FusedCoilGPU.jl:70-133) and only excitation blocks store
 * states, each at most max_rf_adc of them.  A mixed plan does the same: its
 * precession blocks, run whole (hom_run) or one at a time
 * (hom_precession_block), keep their states on chip. */
int hom_plan_create(hom_plan **plan, int device, int precision, int64_t n_spins, int64_t n_coils,
                    int64_t n_nodes, int64_t max_adc, int64_t max_rf_adc, double gamma,
                    int coil_combination);

/* The remaining HighOrderGPUWorkspace keyword checks (reference
 * src/simulation/kernels/HighOrderGPU.jl:24-35), before any allocation:
 * in a fused coil mode, groupsize a multiple of 32 in [32, 1024]; the ADC and
 * coil tiles (both 1 in "fused_warp") from {1, 2, 4, 8} and {1, 2, 4}; and
 * groupsize * adc_tile complex values of the precision within the device's
 * shared memory per block (device < 0: the 64 KiB a workgroup may always
 * use).  HOM_ERR_ARGUMENT with the reference's message otherwise.  The
 * knobs select CUDA launch shapes in the reference; this library's fused
 * kernel has one shape (256 spins, 16-sample tiles), so they are validated
 * and not otherwise used. */
int hom_check_workspace_options(int device, int precision, int coil_combination, int64_t groupsize,
                                int64_t adc_tile, int64_t coil_tile);
int hom_plan_destroy(hom_plan *plan);

/* Device residency of HO_Phantom fields (reference src/datatypes/HO_Phantom.jl:47-71;
 * upload step src/simulation/Simulate.jl:151-157).  csm is Ns x Nc column-major
 * complex double with column stride ld (>= Ns). */
int hom_upload_phantom(hom_plan *plan, const double *x, const double *y, const double *z,
                       const double *rho, const double *T1, const double *T2, const double *dw);
int hom_upload_csm(hom_plan *plan, const double *csm, int64_t ld);
int hom_upload_sequence(hom_plan *plan, const hom_sequence_host *seq);

/* Spin state Mag(xy, z).  reset: Mxy = 0, Mz = rho (reference
 * src/simulation/Bloch/BlochHighOrderSimulatitonMethod.jl:9-15).  get/set
 * copy Ns complex double Mxy and Ns double Mz (synchronous). */
int hom_reset_state(hom_plan *plan);
int hom_set_state(hom_plan *plan, const double *mxy, const double *mz);
int hom_get_state(hom_plan *plan, double *mxy, double *mz);

/* Per-block drop-ins for run_high_order_precession! / run_high_order_excitation!
 * (reference src/simulation/kernels/HighOrderGPU.jl:124-140 and :223-233):
 * evolve the state over nodes [first, last] and write the block's
 * Nadc x Nc signal (overwrite, transpose CSM, no conjugation) into device
 * memory out_dev of the plan's signal type with column stride out_ld.
 * out_dev may be NULL when the block has no ADC sample.  f32/f64 plans:
 * stream-ordered, the call returns before the device finishes.  Mixed plans
 * run a precession block as one mixed segment (out_dev is complex double);
 * the sequence tables are built on the host for that block, so the call may
 * wait for the plan's stream and allocate device memory for the tables and
 * the segment buffers (HOM_ERR_MEMORY on failure). */
int hom_precession_block(hom_plan *plan, int64_t first, int64_t last, int flags, void *out_dev,
                         int64_t out_ld);
int hom_excitation_block(hom_plan *plan, int64_t first, int64_t last, int flags, void *out_dev,
                         int64_t out_ld);

/* reduce_high_order_signal! (reference src/simulation/kernels/CoilReductionGPU.jl:79-100):
 * deterministic two-stage coil weighting and spin reduction of a device state
 * matrix (Ns x n_adc, column stride ld_states, plan's complex type) with the
 * plan's CSM into out_dev (n_adc x Nc, column stride out_ld). */
int hom_reduce_signal(hom_plan *plan, const void *states_dev, int64_t ld_states, int64_t n_adc,
                      void *out_dev, int64_t out_ld);

/* No counterpart in the reference.  hom_reduce_signal for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same kernels and the same
 * result, bit for bit, with the stage-1 partials in a library-owned device
 * array of HOM_REDUCE_SCRATCH_BYTES instead of a plan buffer.  Each device
 * holds one such array from library load until the process ends; it is
 * looked up when a plan is created on that device.  ceil(Ns / 2048) * n_adc *
 * Nc complex values of the plan's type must fit (HOM_ERR_DIMENSION otherwise:
 * at 1,048,576 spins and FP64, n_adc * Nc <= 4096).  The call allocates
 * nothing, records no kernel timers and is stream-ordered.  Every call on one
 * device shares the array, so calls that overlap on one device (two plans'
 * streams, or two threads) race: serialize them. */
#define HOM_REDUCE_SCRATCH_BYTES (32 * 1024 * 1024)
int hom_reduce_signal_noalloc(hom_plan *plan, const void *states_dev, int64_t ld_states,
                              int64_t n_adc, void *out_dev, int64_t out_ld);

/* Whole block loop (reference src/simulation/SimulatorCore.jl:96-152) over
 * the given parts, from the current state, writing sum(ADC) x Nc rows into the
 * plan's signal buffer.  HOM_F32/HOM_F64 plans run the per-block kernels;
 * HOM_MIXED plans fuse consecutive precession blocks into segments. */
int hom_run(hom_plan *plan, const int64_t *part_first, const int64_t *part_last,
            const uint8_t *part_is_rf, int64_t n_parts, int flags);
int hom_synchronize(hom_plan *plan);
int hom_signal_rows(hom_plan *plan, int64_t *rows);
/* Copies the signal buffer (rows x Nc, column stride ld on the host). */
int hom_download_signal(hom_plan *plan, void *host_out, int64_t ld);
/* Element size in bytes of one complex signal value (8 or 16). */
int hom_signal_element_size(hom_plan *plan, int64_t *bytes);

/* Device buffers for hosts without a GPU array library of their own. */
int hom_buffer_alloc(hom_plan *plan, int64_t bytes, void **dev);
int hom_buffer_free(hom_plan *plan, void *dev);
int hom_buffer_upload(hom_plan *plan, void *dev, const void *host, int64_t bytes);
int hom_buffer_download(hom_plan *plan, void *host, const void *dev, int64_t bytes);

/* Per-kernel timing with HIP events on the stream each kernel runs on (off by
 * default). */
int hom_plan_set_timing(hom_plan *plan, int enable);
int hom_plan_kernel_time(hom_plan *plan, int kernel_id, double *total_ms, int64_t *launches);

/* Spin tile of the mixed kernels (no counterpart in the reference, whose CUDA
 * launch shapes are fixed): 16-spin groups per workgroup work item, a power
 * of two in [16, 4096].  Mixed plans only; takes effect from the next hom_run
 * / hom_precession_block.
 *   hom_plan_set_spin_tile       fixes the tile for every segment.
 *   hom_plan_set_spin_tile_auto  makes `groups` an upper bound: per segment
 *                                the library takes the largest tile whose
 *                                work items fill the GPU's workgroup slots
 *                                in whole waves (smaller tiles only for
 *                                small phantoms or short segments).  This is
 *                                the default, with bound 256.
 * hom_plan_get_spin_tile reads the set value; hom_plan_last_spin_tile the
 * tile the longest segment of the last mixed run used.  The shipped tuner
 * (highordermrisim.jl_amd/tune.py) measures the bound per coil-block count
 * and field-order class and persists it. */
int hom_plan_set_spin_tile(hom_plan *plan, int groups);
int hom_plan_set_spin_tile_auto(hom_plan *plan, int groups);
int hom_plan_get_spin_tile(const hom_plan *plan, int *groups);
int hom_plan_last_spin_tile(const hom_plan *plan, int *groups);

/* Diagnostics of the last hom_run with precision "mixed" (no counterpart in
 * the reference): the number of free-precession segments that acquired
 * samples, and how many of them ran the all-FP64 phase body because their
 * higher-order phase offsets within a 256-sample anchor block exceed half a
 * revolution (slower, same accuracy). */
int hom_plan_mixed_stats(const hom_plan *plan, int64_t *segments, int64_t *fp64_segments);

/* Encoding coefficients of a discretized sequence: get_kspace / grad2kspace
 * (reference src/datatypes/HO_Sequence.jl:221-287) on the node grid, host
 * only (no plan, no device).  Outputs are column-major with n_nodes - 1 rows
 * (row i = end of interval i) or one row per ADC node: k_nominal (3 columns,
 * Gx Gy Gz, cycles/m), k_dfc (16 columns, h0..h15, cycles/m^order), and their
 * ADC rows; any output may be NULL.  skip_rf (one flag per RF event, or NULL)
 * continues the integral through that event instead of resetting
 * (excitation, flip <= 90.01 degrees) or negating it (refocusing).  Needs at
 * least two nodes (ArgumentError otherwise, as the reference). */
int hom_kspace(const hom_sequence_host *seq, double gamma, const uint8_t *skip_rf, int64_t n_skip,
               double *k_nominal, double *k_nominal_adc, double *k_dfc, double *k_dfc_adc);
/* RF events (maximal runs of excitation intervals) and their flip angles in
 * degrees (first `capacity` written), as classified by hom_kspace. */
int hom_kspace_rf_events(const hom_sequence_host *seq, double gamma, int64_t *n_events,
                         double *flip_deg, int64_t capacity);

/* Sequence discretization (reference src/datatypes/HO_DiscreteSequence.jl:93-179,
 * with KomaMRIBase's sampling replaced by this library's rules; host only).
 * A waveform is piecewise linear through its knots (times relative to the
 * block start, nondecreasing; a repeated time is a jump) and zero outside
 * them.  B1 is nonzero only inside the RF window [rf_t0, rf_t1). */
typedef struct hom_wave {
    int64_t n;
    const double *t, *a;
} hom_wave;

typedef struct hom_block {
    double duration;                  /* s */
    hom_wave gx, gy, gz;              /* nominal gradients, T/m */
    hom_wave rf_re, rf_im;            /* B1, T */
    double rf_t0, rf_t1;              /* RF window (rf_t1 <= rf_t0: no RF) */
    double rf_df, rf_phase;           /* Hz, rad, inside the RF window */
    int64_t adc_n;
    const double *adc_t;              /* ADC sample times in (0, duration] */
    hom_wave h[16];                   /* field coefficients h0..h15 */
} hom_block;

/* Samples every waveform on one node grid: block edges, knots, ADC times and
 * RF window edges, filled to steps <= dt_max (<= dt_rf inside RF); a value
 * jump keeps both sides at the same time (zero-length interval); blocks share
 * their boundary node when every value agrees.  With nominal_in_h the nominal
 * gradients are added to h1..h3 (measured first-order fields).  The result
 * is held in *out (n_nodes nodes) until hom_discrete_free. */
typedef struct hom_discrete hom_discrete;
int hom_discretize(const hom_block *blocks, int64_t n_blocks, double dt_max, double dt_rf,
                   int nominal_in_h, hom_discrete **out, int64_t *n_nodes);
/* Copies the node arrays (any may be NULL): n values each, b1 interleaved
 * complex (2 n), excitation n - 1, h 16 rows of n (row j contiguous). */
int hom_discrete_copy(const hom_discrete *d, double *t, double *gx, double *gy, double *gz,
                      double *b1, double *df, double *psi, uint8_t *adc, uint8_t *excitation,
                      double *h);
int hom_discrete_free(hom_discrete *d);

#ifdef __cplusplus
}
#endif
#endif
