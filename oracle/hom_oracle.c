/*
 * hom_oracle.c -- CPU f64 restatement of HighOrderMRISim.jl's BlochHighOrder
 * path.  TEST INFRASTRUCTURE ONLY: this is the parity checker for the HIP
 * implementation and the "port" CPU baseline.  It is never part of the
 * product; tests/, the smoke entry point and bench.py's CPU leg are its only
 * callers.
 *
 * Followed sources (paths relative to the reference checkout):
 *   src/simulation/Bloch/BlochHighOrderSimulatitonMethod.jl:43-94   precession
 *   src/simulation/Bloch/BlochHighOrderSimulatitonMethod.jl:100-134  excitation
 *   src/simulation/Bloch/BlochHighOrderSimulatitonMethod.jl:17-20    coil combine
 *   src/simulation/Bloch/BlochHighOrderSimulatitonMethod.jl:9-15     initial state
 *   src/simulation/SimulatorCore.jl:23-35, 96-152                    block loop
 *   src/simulation/kernels/HighOrderGPU.jl:181-192                   spinor update
 *
 * Pinning (each checked in tests/test_oracle_*.py): closed-form per-channel
 * phase, zeroth-order readout phase, 90-degree hard pulse with transpose CSM,
 * ideal-coil and CSM linearity, RF ignoring high-order fields, trapezoid
 * exactness on linear ramps, T1/T2 closed forms, block-length invariance, and
 * a Rodrigues-rotation brute force of the spinor update.
 */
#include "hom_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO_PI 6.283185307179586476925286766559

int64_t hom_oracle_block_adc(const hom_oracle_sequence *seq, int64_t first, int64_t last)
{
    int64_t n = 0;
    for (int64_t i = first + 1; i <= last; ++i) n += seq->adc[i] ? 1 : 0;
    return n;
}

/* Bz of one spin at one node, in the evaluation order of
 * BlochHighOrderSimulatitonMethod.jl:54-75. */
static double field_at(const hom_oracle_sequence *seq, int64_t n, double x, double y, double z,
                       double dw_term, int flags)
{
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = (x2 + y2) + z2;
    double bz0 = 0.0, bz1, bz2 = 0.0, bz3 = 0.0;
    if (flags & HOM_ORACLE_HO0) bz0 = seq->h[0][n];
    if (flags & HOM_ORACLE_HO1)
        bz1 = (x * seq->h[1][n] + y * seq->h[2][n]) + z * seq->h[3][n];
    else
        bz1 = (x * seq->gx[n] + y * seq->gy[n]) + z * seq->gz[n];
    if (flags & HOM_ORACLE_HO2) {
        bz2 = (x * y) * seq->h[4][n];
        bz2 += (z * y) * seq->h[5][n];
        bz2 += (3.0 * z2 - r2) * seq->h[6][n];
        bz2 += (x * z) * seq->h[7][n];
        bz2 += (x2 - y2) * seq->h[8][n];
    }
    if (flags & HOM_ORACLE_HO3) {
        bz3 = ((3.0 * y) * x2 - y * y * y) * seq->h[9][n];
        bz3 += ((x * z) * y) * seq->h[10][n];
        bz3 += ((5.0 * z2 - r2) * y) * seq->h[11][n];
        bz3 += (5.0 * (z * z * z) - (3.0 * z) * r2) * seq->h[12][n];
        bz3 += ((5.0 * z2 - r2) * x) * seq->h[13][n];
        bz3 += (x2 * z - y2 * z) * seq->h[14][n];
        bz3 += (x * x * x - (3.0 * x) * y2) * seq->h[15][n];
    }
    double bz = ((bz0 + bz1) + bz2) + bz3;
    if (flags & HOM_ORACLE_DW_PRECESSION) bz += dw_term;
    return bz;
}

int hom_oracle_precession(const hom_oracle_phantom *ph, const hom_oracle_sequence *seq,
                          int64_t first, int64_t last, int flags, double gamma,
                          int64_t s0, int64_t s1,
                          double *mxy, double *mz, double *sig, int64_t ld_sig)
{
    if (first < 0 || last < first || last >= seq->n_nodes) return -1;
    const int64_t nadc = hom_oracle_block_adc(seq, first, last);
    const int64_t nc = ph->n_coils;
    const double k_phase = -TWO_PI * gamma;          /* T(-2pi*gamma) */
    const double k_dw = 1.0 / (TWO_PI * gamma);      /* Delta-w / T(2pi*gamma) */
    double *acc = NULL;
    int64_t *adc_nodes = NULL;
    if (nadc > 0) {
        if (!sig) return -2;
        acc = (double *)calloc((size_t)(2 * nadc * nc), sizeof(double));
        adc_nodes = (int64_t *)malloc((size_t)nadc * sizeof(int64_t));
        int64_t k = 0;
        for (int64_t n = first + 1; n <= last; ++n)
            if (seq->adc[n]) adc_nodes[k++] = n;
    }
    double *cs = (double *)malloc((size_t)(2 * (nc > 0 ? nc : 1)) * sizeof(double));

    for (int64_t s = s0; s < s1; ++s) {
        const double x = ph->x[s], y = ph->y[s], z = ph->z[s];
        const double dw_term = ph->dw[s] * k_dw;
        const double m_re = mxy[2 * s], m_im = mxy[2 * s + 1];
        const double T2 = ph->T2[s];
        /* cumtrapz (KomaMRIBase) and cumsum(dt), both sequential. */
        double b_prev = field_at(seq, first, x, y, z, dw_term, flags);
        double integral = 0.0, tp = 0.0;
        int64_t k = 0;
        if (nadc > 0)
            for (int64_t c = 0; c < nc; ++c) {
                cs[2 * c] = ph->csm[2 * (s + c * ph->ld_csm)];
                cs[2 * c + 1] = ph->csm[2 * (s + c * ph->ld_csm) + 1];
            }
        for (int64_t n = first + 1; n <= last; ++n) {
            const double dt = seq->dt[n - 1];
            const double b_next = field_at(seq, n, x, y, z, dw_term, flags);
            integral += (b_prev + b_next) * (dt / 2.0);
            tp += dt;
            b_prev = b_next;
            if (k < nadc && adc_nodes[k] == n) {
                const double phi = k_phase * integral;
                const double mag = exp(-tp / T2);
                const double e_re = mag * cos(phi), e_im = mag * sin(phi);
                const double st_re = m_re * e_re - m_im * e_im;
                const double st_im = m_re * e_im + m_im * e_re;
                double *row = acc + 2 * k * nc;
                for (int64_t c = 0; c < nc; ++c) {
                    row[2 * c] += st_re * cs[2 * c] - st_im * cs[2 * c + 1];
                    row[2 * c + 1] += st_re * cs[2 * c + 1] + st_im * cs[2 * c];
                }
                ++k;
            }
        }
        /* End of block: Mxy = last column, Mz relaxation with dur = sum(dt). */
        if (last > first) {
            const double phi = k_phase * integral;
            const double mag = exp(-tp / T2);
            const double e_re = mag * cos(phi), e_im = mag * sin(phi);
            mxy[2 * s] = m_re * e_re - m_im * e_im;
            mxy[2 * s + 1] = m_re * e_im + m_im * e_re;
        }
        const double e1 = exp(-tp / ph->T1[s]);
        mz[s] = mz[s] * e1 + ph->rho[s] * (1.0 - e1);
    }
    if (nadc > 0) {
        for (int64_t c = 0; c < nc; ++c)
            for (int64_t k = 0; k < nadc; ++k) {
                sig[2 * (k + c * ld_sig)] = acc[2 * (k * nc + c)];
                sig[2 * (k + c * ld_sig) + 1] = acc[2 * (k * nc + c) + 1];
            }
    }
    free(cs);
    free(acc);
    free(adc_nodes);
    return 0;
}

/* Hard-pulse spinor update of one spin (HighOrderGPU.jl:181-192, the loop
 * form of KomaMRICore's Q/mul! used by BlochHighOrderSimulatitonMethod.jl:122). */
static void rotate(double *m_re, double *m_im, double *mzp, double b1_re, double b1_im,
                   double bz, double dt, double gamma)
{
    double B = sqrt(b1_re * b1_re + b1_im * b1_im + bz * bz);
    if (B == 0.0) B = 2.220446049250313e-16; /* eps(Float64) */
    const double phi = (-TWO_PI * gamma) * (B * dt);
    const double sn = sin(phi / 2.0), cn = cos(phi / 2.0);
    /* alpha = cos - i (Bz/B) sin ; beta = -i (B1/B) sin */
    const double a_re = cn, a_im = -(bz / B) * sn;
    const double nb_re = b1_re / B, nb_im = b1_im / B;
    const double b_re = nb_im * sn, b_im = -nb_re * sn; /* -i * (nb) * sn */
    const double mx = *m_re, my = *m_im, z = *mzp;
    /* conj(alpha) */
    const double ca_re = a_re, ca_im = -a_im;
    /* 2 conj(a) b Mz */
    const double t1_re = 2.0 * (ca_re * b_re - ca_im * b_im) * z;
    const double t1_im = 2.0 * (ca_re * b_im + ca_im * b_re) * z;
    /* conj(a)^2 Mxy */
    const double ca2_re = ca_re * ca_re - ca_im * ca_im, ca2_im = 2.0 * ca_re * ca_im;
    const double t2_re = ca2_re * mx - ca2_im * my, t2_im = ca2_re * my + ca2_im * mx;
    /* beta^2 conj(Mxy) */
    const double b2_re = b_re * b_re - b_im * b_im, b2_im = 2.0 * b_re * b_im;
    const double t3_re = b2_re * mx + b2_im * my, t3_im = b2_im * mx - b2_re * my;
    /* (|a|^2 - |b|^2) Mz - 2 Re(a b conj(Mxy)) */
    const double ab_re = a_re * b_re - a_im * b_im, ab_im = a_re * b_im + a_im * b_re;
    const double re_ab_cm = ab_re * mx + ab_im * my;
    const double na = a_re * a_re + a_im * a_im, nb = b_re * b_re + b_im * b_im;
    *m_re = t1_re + t2_re - t3_re;
    *m_im = t1_im + t2_im - t3_im;
    *mzp = (na - nb) * z - 2.0 * re_ab_cm;
}

int hom_oracle_excitation(const hom_oracle_phantom *ph, const hom_oracle_sequence *seq,
                          int64_t first, int64_t last, int flags, double gamma,
                          int64_t s0, int64_t s1,
                          double *mxy, double *mz, double *sig, int64_t ld_sig)
{
    if (first < 0 || last < first || last >= seq->n_nodes) return -1;
    const int64_t nadc = hom_oracle_block_adc(seq, first, last);
    const int64_t nc = ph->n_coils;
    if (nadc > 0 && !sig) return -2;
    double *acc = nadc > 0 ? (double *)calloc((size_t)(2 * nadc * nc), sizeof(double)) : NULL;
    const double inv_g = 1.0 / gamma, k_dw = 1.0 / (TWO_PI * gamma);
    for (int64_t s = s0; s < s1; ++s) {
        const double x = ph->x[s], y = ph->y[s], z = ph->z[s];
        double m_re = mxy[2 * s], m_im = mxy[2 * s + 1], m_z = mz[s];
        {   /* M.xy .*= cis(-psi[1]) */
            const double p = -seq->psi[first];
            const double c = cos(p), sn = sin(p);
            const double r = m_re * c - m_im * sn, i = m_re * sn + m_im * c;
            m_re = r; m_im = i;
        }
        int64_t k = 0;
        for (int64_t n = first; n < last; ++n) {
            const double dt = seq->dt[n];
            const double dbz = (flags & HOM_ORACLE_DW_EXCITATION)
                                   ? ph->dw[s] * k_dw - seq->df[n] * inv_g
                                   : 0.0 - seq->df[n] * inv_g;
            const double bz = ((seq->gx[n] * x + seq->gy[n] * y) + seq->gz[n] * z) + dbz;
            rotate(&m_re, &m_im, &m_z, seq->b1[2 * n], seq->b1[2 * n + 1], bz, dt, gamma);
            const double e2 = exp(-dt / ph->T2[s]);
            m_re *= e2; m_im *= e2;
            const double e1 = exp(-dt / ph->T1[s]);
            m_z = m_z * e1 + ph->rho[s] * (1.0 - e1);
            if (seq->adc[n + 1]) {
                const double p = seq->psi[n + 1];
                const double c = cos(p), sn = sin(p);
                const double st_re = m_re * c - m_im * sn, st_im = m_re * sn + m_im * c;
                double *row = acc + 2 * k * nc;
                for (int64_t cc = 0; cc < nc; ++cc) {
                    const double *w = ph->csm + 2 * (s + cc * ph->ld_csm);
                    row[2 * cc] += st_re * w[0] - st_im * w[1];
                    row[2 * cc + 1] += st_re * w[1] + st_im * w[0];
                }
                ++k;
            }
        }
        {   /* M.xy .*= cis(psi[end]) */
            const double p = seq->psi[last];
            const double c = cos(p), sn = sin(p);
            const double r = m_re * c - m_im * sn, i = m_re * sn + m_im * c;
            m_re = r; m_im = i;
        }
        mxy[2 * s] = m_re; mxy[2 * s + 1] = m_im; mz[s] = m_z;
    }
    if (nadc > 0) {
        for (int64_t c = 0; c < nc; ++c)
            for (int64_t k = 0; k < nadc; ++k) {
                sig[2 * (k + c * ld_sig)] = acc[2 * (k * nc + c)];
                sig[2 * (k + c * ld_sig) + 1] = acc[2 * (k * nc + c) + 1];
            }
    }
    free(acc);
    return 0;
}

void hom_oracle_combine(const double *states, int64_t ld_states, int64_t n_spins,
                        int64_t n_adc, const double *csm, int64_t ld_csm, int64_t n_coils,
                        double *sig, int64_t ld_sig)
{
    for (int64_t c = 0; c < n_coils; ++c)
        for (int64_t k = 0; k < n_adc; ++k) {
            double re = 0.0, im = 0.0;
            for (int64_t s = 0; s < n_spins; ++s) {
                const double *a = states + 2 * (s + k * ld_states);
                const double *w = csm + 2 * (s + c * ld_csm);
                re += a[0] * w[0] - a[1] * w[1];
                im += a[0] * w[1] + a[1] * w[0];
            }
            sig[2 * (k + c * ld_sig)] = re;
            sig[2 * (k + c * ld_sig) + 1] = im;
        }
}

static int run_partition(const hom_oracle_phantom *ph, const hom_oracle_sequence *seq,
                         const int64_t *pf, const int64_t *pl, const uint8_t *rf, int64_t np,
                         int flags, double gamma, int64_t s0, int64_t s1,
                         double *mxy, double *mz, double *sig, int64_t ld_sig)
{
    int64_t row = 0;
    for (int64_t b = 0; b < np; ++b) {
        const int64_t nadc = hom_oracle_block_adc(seq, pf[b], pl[b]);
        double *out = nadc > 0 ? sig + 2 * row : NULL;
        int rc = rf[b] ? hom_oracle_excitation(ph, seq, pf[b], pl[b], flags, gamma, s0, s1,
                                               mxy, mz, out, ld_sig)
                       : hom_oracle_precession(ph, seq, pf[b], pl[b], flags, gamma, s0, s1,
                                               mxy, mz, out, ld_sig);
        if (rc) return rc;
        row += nadc;
    }
    return 0;
}

int hom_oracle_run(const hom_oracle_phantom *ph, const hom_oracle_sequence *seq,
                   const int64_t *part_first, const int64_t *part_last,
                   const uint8_t *part_is_rf, int64_t n_parts, int flags, double gamma,
                   double *mxy, double *mz, double *sig, int64_t ld_sig, int n_threads)
{
    int64_t total = 0;
    for (int64_t b = 0; b < n_parts; ++b) total += hom_oracle_block_adc(seq, part_first[b], part_last[b]);
    const int64_t nc = ph->n_coils, ns = ph->n_spins;
    if (n_threads < 1) n_threads = 1;
    if (n_threads > ns) n_threads = (int)(ns > 0 ? ns : 1);
    if (n_threads == 1)
        return run_partition(ph, seq, part_first, part_last, part_is_rf, n_parts, flags, gamma,
                             0, ns, mxy, mz, sig, ld_sig);
    /* Private (SigmaADC x Nc) buffer per contiguous spin partition, summed in
     * partition order: Simulate.jl:150 (sig(..., Nthreads)) and :185. */
    const size_t per = (size_t)(2 * total * nc);
    double *priv = (double *)calloc(per * (size_t)n_threads + 1, sizeof(double));
    if (!priv) return -3;
    int err = 0;
#pragma omp parallel for schedule(static, 1) num_threads(n_threads) reduction(| : err)
    for (int i = 0; i < n_threads; ++i) {
        const int64_t s0 = (ns * i) / n_threads, s1 = (ns * (i + 1)) / n_threads;
        err |= run_partition(ph, seq, part_first, part_last, part_is_rf, n_parts, flags, gamma,
                             s0, s1, mxy, mz, priv + per * (size_t)i, total);
    }
    if (!err) {
        for (int64_t c = 0; c < nc; ++c)
            for (int64_t k = 0; k < total; ++k) {
                double re = 0.0, im = 0.0;
                for (int i = 0; i < n_threads; ++i) {
                    re += priv[per * (size_t)i + 2 * (k + c * total)];
                    im += priv[per * (size_t)i + 2 * (k + c * total) + 1];
                }
                sig[2 * (k + c * ld_sig)] = re;
                sig[2 * (k + c * ld_sig) + 1] = im;
            }
    }
    free(priv);
    return err;
}

int64_t hom_oracle_partition(const uint8_t *excitation, int64_t n_intervals,
                             int64_t max_len, int64_t max_rf_len,
                             int64_t *part_first, int64_t *part_last, uint8_t *part_is_rf,
                             int64_t cap)
{
    int64_t count = 0, i = 0;
    if (max_len < 1) max_len = 1;
    if (max_rf_len < 1) max_rf_len = 1;
    while (i < n_intervals) {
        const uint8_t f = excitation[i] ? 1 : 0;
        int64_t j = i;
        while (j < n_intervals && (excitation[j] ? 1 : 0) == f) ++j;
        const int64_t lim = f ? max_rf_len : max_len;
        for (int64_t a = i; a < j; a += lim) {
            const int64_t b = (a + lim < j) ? a + lim : j;
            if (count < cap) {
                part_first[count] = a;
                part_last[count] = b;
                part_is_rf[count] = f;
            }
            ++count;
        }
        i = j;
    }
    return count;
}
