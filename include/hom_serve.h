/*
 * hom_serve.h -- continuous-batching simulation server of libhom.so.
 *
 * A server owns one device, one HIP stream, a device arena (the memory
 * budget that requests are admitted against) and a pinned host arena for
 * results.  Phantoms, sensitivity maps and sequences are uploaded once and
 * stay resident in HBM; a request names resident objects plus a spin range,
 * a coil count, a node range and BlochHighOrder flags, and is simulated with
 * the mixed-precision path (FP64 phase, split-FP16 MFMA coil GEMM, FP64
 * signal).
 *
 * Scheduling: the server advances in steps.  Each step admits queued
 * requests in arrival order while the arena has room and a slot is free,
 * then runs one chunk (whole 256-sample anchor blocks, up to the per-request
 * step budget) of every active request in one set of persistent batched
 * launches, replayed from a captured hipGraph per set of kernel classes.
 * Requests that arrive while others are on the device join at the next step
 * boundary; finished requests leave and free their memory.  A request's
 * result does not depend on what it was batched with: every reduction is
 * private to the request and runs in a fixed order.
 *
 * This has no counterpart in the reference, which simulates one
 * (phantom, sequence) per call (src/simulation/Simulate.jl:52-135) and keeps
 * inputs resident only per PreparedSimulation
 * (src/simulation/PreparedSimulation.jl:2-287); the per-request arithmetic is
 * that of hom_run with HOM_MIXED (hom_abi.h).
 *
 * Thread safety: every entry point may be called from any thread; the
 * server serializes them internally.  With threaded = 1 a worker thread runs
 * steps whenever work is queued; otherwise the caller drives hom_server_step.
 * hom_server_destroy is the exception: no other call on the server may be in
 * progress or follow it (it stops the worker, abandons unfinished requests
 * and frees everything).
 *
 * Errors: a device error, or a host allocation failure inside a step, stops
 * the server.  Every request that has not finished (active or queued) turns
 * HOM_REQ_FAILED, waiters return, and later submissions and steps fail with
 * HOM_ERR_DEVICE; hom_server_last_error holds the cause.
 */
#ifndef HOM_SERVE_H
#define HOM_SERVE_H

#include <stdint.h>

#include "hom_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hom_server hom_server;

#define HOM_SERVE_MAX_ACTIVE 64

typedef struct hom_server_config {
    int device;
    int64_t device_bytes;  /* device arena (admission budget); 0: 8 GiB */
    int64_t host_bytes;    /* pinned result arena; 0: 1 GiB */
    int max_active;        /* requests per step, 1..HOM_SERVE_MAX_ACTIVE; 0: 64 */
    int64_t step_budget;   /* spin-samples per request per step; 0: 2^33 */
    int use_graphs;        /* 1: replay captured step graphs; 0: launch directly */
    int gpt;               /* 16-spin groups per spin tile: 64, 128, 256 or 512; 0: 256 */
    int main_grid;         /* persistent workgroups of the fused kernels; 0: 3 per CU */
    int threaded;          /* 1: a worker thread runs steps */
} hom_server_config;

/* Request states. */
enum hom_request_state {
    HOM_REQ_QUEUED = 0,
    HOM_REQ_ACTIVE = 1,
    HOM_REQ_DONE = 2,
    HOM_REQ_FAILED = 3
};

/* Kinds of resident objects for hom_server_release. */
enum hom_resident_kind { HOM_RES_PHANTOM = 0, HOM_RES_CSM = 1, HOM_RES_SEQUENCE = 2 };

typedef struct hom_request {
    int64_t phantom, spin_first, spin_count; /* spins [spin_first, spin_first + spin_count) */
    int64_t csm, coil_count;                 /* the first coil_count coils of a resident CSM */
    int64_t sequence, node_first, node_last; /* inclusive node range of a resident sequence */
    int flags;                               /* hom_flags */
} hom_request;

typedef struct hom_server_stats {
    int64_t submitted, completed, failed, queued, active;
    int64_t steps, graph_launches, graphs_captured;
    int64_t max_active_seen;
    int64_t device_bytes_used, device_bytes_peak;
    double busy_seconds;          /* wall time inside steps */
    double spin_samples;          /* completed spin-samples */
} hom_server_stats;

int hom_server_create(hom_server **server, const hom_server_config *config);
int hom_server_destroy(hom_server *server);
/* Last error of this server; with server == NULL, the calling thread's last
 * error from hom_server_create. */
const char *hom_server_last_error(const hom_server *server);
/* The configuration in use (defaults resolved). */
int hom_server_get_config(hom_server *server, hom_server_config *out);

/* Resident store.  Phantom arrays have n_spins doubles each; csm is
 * n_spins x n_coils column-major complex double with column stride ld; gamma
 * (Hz/T) travels with the sequence.  Ids are positive. */
int hom_server_put_phantom(hom_server *server, int64_t n_spins, const double *x, const double *y,
                           const double *z, const double *rho, const double *T1, const double *T2,
                           const double *dw, int64_t *id);
int hom_server_put_csm(hom_server *server, int64_t phantom, int64_t n_coils, const double *csm,
                       int64_t ld, int64_t *id);
int hom_server_put_sequence(hom_server *server, const hom_sequence_host *seq, double gamma,
                            int64_t *id);
/* Frees a resident object once no queued or active request uses it. */
int hom_server_release(hom_server *server, int kind, int64_t id);

/* Queues a request; validation errors are returned here.  A request whose
 * footprint exceeds the whole device arena is rejected with HOM_ERR_MEMORY. */
int hom_server_submit(hom_server *server, const hom_request *req, int64_t *request_id);
/* Device bytes a request would hold while active. */
int hom_server_footprint(hom_server *server, const hom_request *req, int64_t *bytes);
/* One scheduler step (admit, run one chunk of every active request, retire);
 * *worked = 0 when there was nothing to do.  Not allowed with threaded = 1. */
int hom_server_step(hom_server *server, int *worked);
/* Runs steps (or waits for the worker) until nothing is queued or active. */
int hom_server_drain(hom_server *server);
/* Waits up to timeout_s (< 0: no limit) for a request; *state gets its state.
 * HOM_ERR_ARGUMENT when the id is unknown, or when another thread fetched the
 * request's result (which forgets it) during the wait. */
int hom_server_wait(hom_server *server, int64_t request_id, double timeout_s, int *state);
int hom_server_status(hom_server *server, int64_t request_id, int *state);
/* Rows (total ADC samples) and columns (coils) of a request's signal. */
int hom_server_result_shape(hom_server *server, int64_t request_id, int64_t *rows, int64_t *cols);
/* Copies a finished request's signal (rows x cols complex double, column
 * stride ld >= rows) and forgets the request. */
int hom_server_result(hom_server *server, int64_t request_id, double *out, int64_t ld);
/* Seconds on a steady clock: submission, admission, completion. */
int hom_server_times(hom_server *server, int64_t request_id, double *t_submit, double *t_admit,
                     double *t_done);
int hom_server_get_stats(hom_server *server, hom_server_stats *out);

/* Autotuner: times the fused kernels of one step of a probe request for each
 * (gpt, main_grid) candidate pair (n_gpt x n_grid, NULL lists use built-in
 * candidates), keeps the fastest in the server configuration and returns it.
 * Needs an idle server (nothing queued or active). */
int hom_server_autotune(hom_server *server, const hom_request *probe, const int *gpt_list,
                        int n_gpt, const int *grid_list, int n_grid, int *best_gpt,
                        int *best_grid, double *best_seconds);

#ifdef __cplusplus
}
#endif

#endif /* HOM_SERVE_H */
