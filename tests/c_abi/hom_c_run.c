/* hom_c_run -- drives libhom.so through its C ABI alone (include/hom_abi.h):
 * plan creation, phantom / CSM / sequence upload, the whole block loop
 * (hom_run) and the signal download, as a C host that links the library
 * would.  Test driver: reads one problem from a binary file and writes the
 * signal to another.
 *
 *   hom_c_run IN OUT [SPIN_TILE]
 *
 * IN (native-endian): int64 precision, coil_mode, flags, ns, nc, nn,
 * max_adc, max_rf_adc, n_parts; double gamma; double x, y, z, rho, T1, T2, dw
 * [ns each]; double csm [2 ns nc, column-major, interleaved]; double t [nn],
 * dt [nn - 1], gx, gy, gz [nn], b1 [2 nn], df, psi [nn]; uint8 adc [nn],
 * excitation [nn - 1]; double h [16 nn, term j at j nn]; int64 part_first,
 * part_last [n_parts]; uint8 part_is_rf [n_parts].
 * OUT: int64 rows, int64 coils, then double signal [2 rows coils,
 * column-major, interleaved].  Exit status 0, or the failing call's status
 * with hom_last_error() on stderr. */
#include "hom_abi.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void *take(FILE *f, size_t bytes) {
    void *p = malloc(bytes ? bytes : 1);
    if (!p || fread(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "hom_c_run: short input\n");
        exit(90);
    }
    return p;
}

#define CHECK(call)                                                                   \
    do {                                                                              \
        int st_ = (call);                                                             \
        if (st_ != HOM_OK) {                                                          \
            fprintf(stderr, "hom_c_run: %s -> %d: %s\n", #call, st_, hom_last_error(plan)); \
            return st_;                                                               \
        }                                                                             \
    } while (0)

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: hom_c_run IN OUT [SPIN_TILE]\n");
        return 91;
    }
    FILE *in = fopen(argv[1], "rb");
    if (!in) {
        perror(argv[1]);
        return 92;
    }
    int64_t *hdr = take(in, 9 * sizeof(int64_t));
    const int64_t precision = hdr[0], coil_mode = hdr[1], flags = hdr[2], ns = hdr[3], nc = hdr[4],
                  nn = hdr[5], max_adc = hdr[6], max_rf_adc = hdr[7], np = hdr[8];
    double *gamma = take(in, sizeof(double));
    double *sp[7];
    for (int i = 0; i < 7; ++i) sp[i] = take(in, (size_t)ns * sizeof(double));
    double *csm = take(in, (size_t)(2 * ns * nc) * sizeof(double));
    double *t = take(in, (size_t)nn * sizeof(double));
    double *dt = take(in, (size_t)(nn - 1) * sizeof(double));
    double *gx = take(in, (size_t)nn * sizeof(double));
    double *gy = take(in, (size_t)nn * sizeof(double));
    double *gz = take(in, (size_t)nn * sizeof(double));
    double *b1 = take(in, (size_t)(2 * nn) * sizeof(double));
    double *df = take(in, (size_t)nn * sizeof(double));
    double *psi = take(in, (size_t)nn * sizeof(double));
    uint8_t *adc = take(in, (size_t)nn);
    uint8_t *exc = take(in, (size_t)(nn - 1));
    double *h = take(in, (size_t)(16 * nn) * sizeof(double));
    int64_t *pf = take(in, (size_t)np * sizeof(int64_t));
    int64_t *pl = take(in, (size_t)np * sizeof(int64_t));
    uint8_t *rf = take(in, (size_t)np);
    fclose(in);

    hom_plan *plan = NULL;
    if (hom_plan_create(&plan, 0, (int)precision, ns, nc, nn, max_adc, max_rf_adc, *gamma,
                        (int)coil_mode) != HOM_OK) {
        fprintf(stderr, "hom_c_run: hom_plan_create: %s\n", hom_last_error(NULL));
        return 93;
    }
    hom_sequence_host seq;
    seq.n_nodes = nn;
    seq.t = t;
    seq.dt = dt;
    seq.gx = gx;
    seq.gy = gy;
    seq.gz = gz;
    seq.b1 = b1;
    seq.df = df;
    seq.psi = psi;
    seq.adc = adc;
    seq.excitation = exc;
    for (int j = 0; j < 16; ++j) seq.h[j] = h + (size_t)j * (size_t)nn;
    CHECK(hom_upload_phantom(plan, sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], sp[6]));
    CHECK(hom_upload_csm(plan, csm, ns));
    CHECK(hom_upload_sequence(plan, &seq));
    if (argc > 3) CHECK(hom_plan_set_spin_tile(plan, atoi(argv[3])));
    CHECK(hom_reset_state(plan));
    CHECK(hom_run(plan, pf, pl, rf, np, (int)flags));
    int64_t rows = 0, esize = 0;
    CHECK(hom_signal_rows(plan, &rows));
    CHECK(hom_signal_element_size(plan, &esize));
    const size_t n = (size_t)(rows * nc);
    void *sig = malloc(n * (size_t)esize + 1);
    double *out = malloc(2 * n * sizeof(double) + 1);
    if (!sig || !out) return 94;
    CHECK(hom_download_signal(plan, sig, rows > 0 ? rows : 1));
    for (size_t i = 0; i < 2 * n; ++i)
        out[i] = esize == 8 ? (double)((const float *)sig)[i] : ((const double *)sig)[i];
    CHECK(hom_plan_destroy(plan));
    FILE *fo = fopen(argv[2], "wb");
    if (!fo) {
        perror(argv[2]);
        return 95;
    }
    const int64_t shape[2] = {rows, nc};
    fwrite(shape, sizeof(int64_t), 2, fo);
    fwrite(out, sizeof(double), 2 * n, fo);
    fclose(fo);
    return 0;
}
