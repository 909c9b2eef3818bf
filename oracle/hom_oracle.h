/*
 * hom_oracle.h -- CPU f64 restatement of the BlochHighOrder precession,
 * excitation and coil-reception path of HighOrderMRISim.jl.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in highordermrisim.jl_amd/.  Only tests/, the smoke entry
 * point and bench.py's CPU-baseline leg may load it.  The product never links
 * or calls it.
 *
 * Layout conventions (Julia column-major, 0-based node indices here):
 *   - complex values are interleaved {re, im} doubles;
 *   - csm is Ns x Nc, column c starts at csm + 2*c*ld_csm;
 *   - signal blocks are Nadc x Nc, column c starts at sig + 2*c*ld_sig;
 *   - a block is the inclusive node range [first, last]; adjacent blocks share
 *     a node and the ADC flag of that shared node belongs to the earlier block
 *     (src/simulation/SimulatorCore.jl:115-117).
 */
#ifndef HOM_ORACLE_H
#define HOM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field-order / off-resonance flags (src/simulation/BlochHighOrder.jl:21-43). */
enum {
    HOM_ORACLE_HO0 = 1 << 0,
    HOM_ORACLE_HO1 = 1 << 1,
    HOM_ORACLE_HO2 = 1 << 2,
    HOM_ORACLE_HO3 = 1 << 3,
    HOM_ORACLE_DW_EXCITATION = 1 << 4,
    HOM_ORACLE_DW_PRECESSION = 1 << 5
};

typedef struct hom_oracle_phantom {
    int64_t n_spins;
    int64_t n_coils;
    const double *x, *y, *z, *rho, *T1, *T2, *dw;
    const double *csm;   /* interleaved complex, Ns x Nc, column stride ld_csm */
    int64_t ld_csm;
} hom_oracle_phantom;

typedef struct hom_oracle_sequence {
    int64_t n_nodes;
    const double *t;            /* n */
    const double *dt;           /* n-1 */
    const double *gx, *gy, *gz; /* n, T/m */
    const double *b1;           /* n, complex interleaved, T */
    const double *df;           /* n, Hz */
    const double *psi;          /* n, rad */
    const uint8_t *adc;         /* n */
    const uint8_t *excitation;  /* n-1 */
    const double *h[16];        /* 16 coefficient waveforms, n each */
} hom_oracle_sequence;

/* Number of ADC samples owned by block [first, last]: ADC[first+1 .. last]. */
int64_t hom_oracle_block_adc(const hom_oracle_sequence *seq, int64_t first, int64_t last);

/* Free-precession block (BlochHighOrderSimulatitonMethod.jl:43-94).
 * Spins [s0, s1) only; writes their contribution to sig (overwrites) and
 * updates mxy/mz for those spins.  sig may be NULL when the block has no ADC. */
int hom_oracle_precession(const hom_oracle_phantom *ph, const hom_oracle_sequence *seq,
                          int64_t first, int64_t last, int flags, double gamma,
                          int64_t s0, int64_t s1,
                          double *mxy, double *mz, double *sig, int64_t ld_sig);

/* Hard-pulse excitation block (BlochHighOrderSimulatitonMethod.jl:100-134). */
int hom_oracle_excitation(const hom_oracle_phantom *ph, const hom_oracle_sequence *seq,
                          int64_t first, int64_t last, int flags, double gamma,
                          int64_t s0, int64_t s1,
                          double *mxy, double *mz, double *sig, int64_t ld_sig);

/* Whole block loop (SimulatorCore.jl:96-152 with Simulate.jl:138-193):
 * runs every part in order, writing each block's ADC rows in sequence into
 * sig (SigmaADC x Nc, column stride ld_sig).  The state arrays hold the
 * initial state on entry and the final state on return.  n_threads > 1 splits
 * the spins into contiguous partitions with private signal buffers summed in
 * partition order afterwards (SimulatorCore.jl:23-35, Simulate.jl:185). */
int hom_oracle_run(const hom_oracle_phantom *ph, const hom_oracle_sequence *seq,
                   const int64_t *part_first, const int64_t *part_last,
                   const uint8_t *part_is_rf, int64_t n_parts, int flags, double gamma,
                   double *mxy, double *mz, double *sig, int64_t ld_sig, int n_threads);

/* sig[k, c] = sum_s states[s, k] * csm[s, c] (transpose, no conjugation;
 * BlochHighOrderSimulatitonMethod.jl:17-20).  states is Ns x Nadc complex,
 * column stride ld_states. */
void hom_oracle_combine(const double *states, int64_t ld_states, int64_t n_spins,
                        int64_t n_adc, const double *csm, int64_t ld_csm, int64_t n_coils,
                        double *sig, int64_t ld_sig);

/* Block partition: maximal runs of equal excitation flag over the intervals,
 * each split into chunks of at most max_len (RF: max_rf_len) intervals.  Writes
 * up to cap parts and returns the total count (call with cap = 0 to size). */
int64_t hom_oracle_partition(const uint8_t *excitation, int64_t n_intervals,
                             int64_t max_len, int64_t max_rf_len,
                             int64_t *part_first, int64_t *part_last, uint8_t *part_is_rf,
                             int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
