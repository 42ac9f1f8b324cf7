/* hpglx.h -- device-pointer entry points of the MI355X kriging library.
 *
 * These take device buffers (for example torch tensors' data_ptr()) and a
 * HIP stream, enqueue the whole pipeline on that stream and return without
 * synchronising.  They extend the reference ABI (include/hpgl.h) with what
 * the reference computes but does not return (kriging variance, neighbour
 * lists), with a mixed-precision solver and with multi-threshold indicator
 * kriging.  Each cites the reference entry point whose work it performs.
 *
 * All entry points return 0 on success and -1 on error; hpglx_last_error()
 * then holds the message.  Calls are stream-ordered: buffers handed to a call
 * may be reused once the call's work on its stream has completed.  The
 * library keeps one workspace per process; a call waits for the previous
 * call's work before reusing it, so concurrent calls on different streams
 * are serialised on the device.
 */
#ifndef HPGL_MI355X_HPGLX_H
#define HPGL_MI355X_HPGLX_H

#include <stddef.h>

#include "hpgl.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { HPGLX_KIND_OK = 0, HPGLX_KIND_SK = 1, HPGLX_KIND_MIK = 2 };
enum { HPGLX_FP64 = 0, HPGLX_MIXED = 1 };
enum { HPGLX_KRIGED = 0, HPGLX_NO_NEIGHBOURS = 1, HPGLX_SINGULAR = 2, HPGLX_INFORMED = 3 };

/* One grid-kriging request.  Arrays are device pointers over the grid in
 * x-fastest order (the reference's F-order property arrays); any output may
 * be null.  `marginal` is a host array. */
typedef struct hpglx_grid_request_t {
    int kind;                  /* HPGLX_KIND_* */
    int precision;             /* HPGLX_FP64 or HPGLX_MIXED (FP32 Cholesky + FP64 refinement) */
    int dims[3];               /* nx, ny, nz */
    int radiuses[3];
    int max_neighbours;        /* 1..32 */
    struct hpgl_cov_params_t cov;
    const float* values;       /* OK/SK: data */
    const unsigned char* mask; /* nonzero = informed */
    const unsigned char* ind;  /* MIK: nind indicator arrays (value 1 = indicator set) */
    const float* zfield;       /* MIK threshold mode: field, indicator_k = (z <= thresholds[k]) */
    const float* thresholds;   /* MIK threshold mode: device array of nind floats */
    int nind;                  /* MIK: number of indicators / thresholds, 1..16 */
    const double* marginal;    /* MIK: host array, marginal P(indicator_k = 1) (the SK mean) */
    int auto_mean;             /* SK: mean of the informed values */
    double mean;               /* SK: explicit mean */
    float* out_val;            /* OK/SK estimate; cells not kriged follow the reference policy */
    unsigned char* out_mask;
    unsigned char* out_class;  /* MIK: nind x N indicator 0/1 (prob > 0.5) */
    float* out_prob;           /* MIK: nind x N probabilities */
    double* out_var;           /* kriging variance; -1 where not kriged, 0 on informed cells */
    double* out_est64;         /* estimate accumulated in FP64 (nind x N for MIK) */
    unsigned char* out_status; /* HPGLX_KRIGED ... HPGLX_INFORMED */
    int order_relations;       /* MIK: correct the CDF (indicator_kriging.h:34-59) */
    int* out_count;            /* neighbour count per cell in bits 0-15; bit 16 set when the cell's
                                  neighbours came from the cluster search (clusterizer.cpp);
                                  -1 on informed OK/SK cells */
    int* out_nboff;            /* N x max_neighbours packed offsets, see hpglx_unpack_offset */
    double pivot_tol;          /* MIXED: re-solve in FP64 when min pivot < tol * C(0); <= 0 -> 1e-3 */
    int refine_steps;          /* MIXED: FP64 residual refinement steps; < 0 -> 1 */
    const float* lvm_mean;     /* SK: locally varying mean, one per cell (replaces auto_mean / mean) */
} hpglx_grid_request_t;

/* The work of hpgl_ordinary_kriging / hpgl_simple_kriging / hpgl_median_ik
 * (api.cpp:253-293, 323-354, 552-601) on device buffers, enqueued on
 * `stream` (a hipStream_t; null = the default stream). */
int hpglx_krige(const hpglx_grid_request_t* req, void* stream);

/* Launch settings for a request, found by timing (no counterpart in the
 * reference): runs the request `reps` times (plus one warm-up) for each
 * candidate -- template entries the neighbour search stages in LDS (1024,
 * 512, 256), persistent blocks per CU of the solve (all that fit, 1), fixed
 * or dynamic dealing of its 16-cell chunks -- and keeps the fastest for
 * later calls with the same parameter set, grid and precision (and in the
 * file named by HPGLX_TUNE_FILE, read at the first call, when set).  The
 * request's outputs hold the last run's results.  Synchronises the stream.
 * chosen (may be null) receives {staged entries (0 = 1024), blocks per CU
 * (0 = all), dynamic}; best_ms the fastest candidate's mean time. */
int hpglx_autotune(const hpglx_grid_request_t* req, void* stream, int reps, int* chosen, float* best_ms);

/* Size of the allocation-free entry point's scratch, per GPU. */
#define HPGLX_NOALLOC_SCRATCH_MIB 128

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the work of hpglx_krige, with the same request and
 * results, calling no hipMalloc or hipFree.  Every device buffer the call
 * needs beyond the caller's inputs and outputs (the parameter set's tables,
 * padded mask, neighbour lists, cluster lists, counters) comes from one
 * static device array of HPGLX_NOALLOC_SCRATCH_MIB MiB that each GPU holds
 * from the library's first such call until the process ends.  A call whose
 * buffers do not fit (the 100x100x50, 24-neighbour grid takes about 80 MiB;
 * the neighbour lists, 4 * cells * max_neighbours bytes, are most of it)
 * returns -1 with a message naming the bytes needed, before any kernel runs.
 * All calls on a GPU share its array, which keeps the last call's contents.
 * The library orders each call after the previous call's work (a stream wait
 * on its completion event), so calls made through it do not overlap; calls
 * that overlap on one GPU anyway (for example one captured into a graph and
 * replayed beside another) race.  Host memory is still used for the tables
 * on the first call with a parameter set. */
int hpglx_krige_noalloc(const hpglx_grid_request_t* req, void* stream);

/* Convenience forms of hpglx_krige. */
int hpglx_ordinary_kriging(const struct hpgl_cov_params_t* cov, const int radiuses[3], int max_neighbours,
                           const int dims[3], const float* values, const unsigned char* mask, float* out_val,
                           unsigned char* out_mask, double* out_var, int precision, void* stream);
int hpglx_simple_kriging(const struct hpgl_cov_params_t* cov, const int radiuses[3], int max_neighbours,
                         const int dims[3], const float* values, const unsigned char* mask, int auto_mean,
                         double mean, float* out_val, unsigned char* out_mask, double* out_var, int precision,
                         void* stream);
/* Multi-threshold median IK on a continuous field: one weight solve per cell,
 * nthr combines, then the order-relation correction of the CDF. */
int hpglx_indicator_cdf(const struct hpgl_cov_params_t* cov, const int radiuses[3], int max_neighbours,
                        const int dims[3], const float* zfield, const unsigned char* mask, int nthr,
                        const float* thresholds, const double* marginal, float* out_cdf, int precision,
                        void* stream);

/* The work of hpgl_lvm_kriging (api.cpp:412-455, lvm_kriging.cpp:16-39):
 * simple kriging with mean_field[cell] as each cell's mean. */
int hpglx_lvm_kriging(const struct hpgl_cov_params_t* cov, const int radiuses[3], int max_neighbours,
                      const int dims[3], const float* values, const unsigned char* mask, const float* mean_field,
                      float* out_val, unsigned char* out_mask, double* out_var, int precision, void* stream);

/* The work of hpgl_indicator_kriging (api.cpp:457-482,
 * indicator_kriging.h:88-161) on device buffers.  `data` holds category
 * indices; `params` is a host array of `ncat` (1..256) per-category models,
 * neighbourhoods and marginal probabilities.  Categories with the same model,
 * radiuses and neighbour limit share one search and one weight solve.
 * out_prob (ncat x N, may be null) receives the probabilities after the
 * order-relation correction; out_cat the first most probable category. */
int hpglx_indicator_kriging(const struct hpgl_ik_params_t* params, int ncat, const int dims[3],
                            const unsigned char* data, const unsigned char* mask, unsigned char* out_cat,
                            unsigned char* out_mask, float* out_prob, int precision, void* stream);

/* Sequential Gaussian simulation on device arrays (values, mask: N each;
 * simulation_mask: N or null; mean_field: N or null, simple kriging with a
 * locally varying mean as hpgl_sgs_lvm_simulation), as hpgl_sgs_simulation;
 * the CDF tables are host arrays.  Returns 0, or -1 with the message in
 * hpgl_get_last_exception_message(). */
int hpglx_sgs_simulation(float* values, unsigned char* mask, const int dims[3], const struct hpgl_sgs_params_t* params,
                         const struct hpgl_non_parametric_cdf_t* cdf, const unsigned char* simulation_mask,
                         const float* mean_field, void* stream);

/* Several realizations of hpglx_sgs_simulation from one conditioning data
 * set (values, mask: device, N; untouched), realization r with seed
 * seeds[r] (host array of nreal, 1..4096): each equals what
 * hpglx_sgs_simulation gives with that seed, written to out_values +
 * r N and out_mask + r N (device, nreal x N).  The realizations advance
 * together, one launch per level index, which fills the GPU where one
 * realization's levels are small.  The reference runs realizations one call
 * at a time (sgs_simulation in geo_bsd); this has no single counterpart. */
int hpglx_sgs_realizations(const float* values, const unsigned char* mask, const int dims[3],
                           const struct hpgl_sgs_params_t* params, const long* seeds, int nreal,
                           const struct hpgl_non_parametric_cdf_t* cdf, const unsigned char* simulation_mask,
                           const float* mean_field, float* out_values, unsigned char* out_mask, void* stream);

/* The host schedule of a sequential simulation, no GPU needed: the cells to
 * simulate (informed = 0 and simulation_mask null or 1) along the
 * reference's random path for `seed` (path_random_generator.cpp:20-94),
 * grouped into levels that run in parallel -- two cells inside each other's
 * +-radiuses box never share a level, and the earlier on the path has the
 * lower one.  Writes, level by level and in path order within a level, each
 * cell, its standard normal deviate (the reference's quantile of the cell's
 * mt19937 draw; NaN where the draw is 0 or 1) and its level.  Returns the
 * number of cells (nothing written when it exceeds capacity), or -1. */
int hpglx_sgs_schedule(const int dims[3], const int radiuses[3], const unsigned char* informed,
                       const unsigned char* simulation_mask, long seed, int* out_cells, double* out_deviates,
                       int* out_level, int capacity);

/* Property files straight to and from device buffers (the work of
 * hpgl_read_inc_file_float / _byte and hpgl_write_inc_file_float / _byte,
 * api.cpp:53-251): the text is parsed and printed on the GPU.  These
 * synchronise `stream` before returning.  gslib = 1 writes the GSLIB form
 * (property_writer.cpp:88-139). */
int hpglx_read_inc_float(const char* filename, float undefined_value, int size, float* d_data, unsigned char* d_mask,
                         void* stream);
int hpglx_read_inc_byte(const char* filename, int undefined_value, int size, unsigned char* d_data,
                        unsigned char* d_mask, const unsigned char* values, int values_count, void* stream);
int hpglx_write_inc_float(const char* filename, const float* d_data, const unsigned char* d_mask, int size,
                          float undefined_value, const char* name, int gslib, void* stream);
int hpglx_write_inc_byte(const char* filename, const unsigned char* d_data, const unsigned char* d_mask, int size,
                         int undefined_value, const char* name, const unsigned char* values, int values_count,
                         int gslib, void* stream);

/* Decodes an entry of out_nboff into the neighbour's offset from the cell. */
void hpglx_unpack_offset(int packed, int* dx, int* dy, int* dz);

/* The ordered search template of a model (covariance_field.h:61-106):
 * returns its length and writes up to `cap` dx,dy,dz triples. */
int hpglx_template(const struct hpgl_cov_params_t* cov, const int radiuses[3], int* out_xyz, int cap);

/* Profiling: when on, the dispatches of each call's search kernel and main
 * solve kernel record start and stop events on the call's stream (no extra
 * packets between kernels).  hpglx_last_timing waits for the last call and
 * returns milliseconds {search kernel, main solve kernel};
 * hpglx_profile_summary waits for every call profiled since profiling was
 * switched on (the newest 4096) and returns their means and their number. */
void hpglx_set_profiling(int on);
int hpglx_last_timing(float* ms2);
int hpglx_profile_summary(float* ms2, int* ncalls);

/* Waits for the last call and reports how many systems the mixed-precision
 * solver re-solved in FP64 and how many cells took the cluster path. */
int hpglx_last_call_stats(unsigned int* fp64_fallbacks, int* cluster_path_cells);

/* Graph replay of repeated grid calls: when on, a device call equal to the
 * previous one (same parameters, geometry and buffers) is captured into a
 * hipGraph once and replayed while the library's workspace is unchanged.
 * Off by default (a replay measured no faster on MI355X); the environment
 * variable HPGLX_GRAPH=1 also turns it on.  No counterpart in the reference. */
void hpglx_set_graph_mode(int on);

/* Precision used by the host-buffer entry points of hpgl.h (default FP64). */
void hpglx_set_default_precision(int precision);

/* Host run of the device introsort (for tests): perm[i] = input position of
 * the element sorted to slot i, ordering keys descending. */
void hpglx_host_introsort_perm(const double* keys, int n, int* perm);
/* The same when only the first kneed slots are wanted: perm[0, kneed) equals
 * the full sort's (the cluster search asks for max_neighbours). */
void hpglx_host_introsort_first(const double* keys, int n, int kneed, int* perm);

/* Leave-one-out ordinary kriging of scattered points (no counterpart in the
 * reference; built on its covariance model, cov_model.h:146-183, and its OK
 * system, my_kriging_weights.h:198-361).  Point i is estimated from its k
 * nearest other points for each of `nsets` parameter sets
 * params[5 p .. 5 p + 4] = (sill, nugget, range_x, range_y, range_z), model
 * 0 spherical, 1 exponential, 2 gaussian.  All arrays are device arrays:
 * xyz n x 3, values n, nbr n x k, est nsets x n (status 0 kriged, 2 singular:
 * estimate = neighbour mean, zero derivative).
 *
 * hpglx_knn: the k nearest other points by Euclidean distance, ties by the
 * lower index, ascending.
 * hpglx_loo_forward: estimates; when `state` is non-null (a device buffer of
 * hpglx_loo_state_size bytes) it keeps the Cholesky factors and weights for
 * the backward pass.
 * hpglx_loo_backward: with grad_est (nsets x n, may be null) the gradient of
 * sum_p,i grad_est[p, i] * est[p, i] w.r.t. each parameter set into
 * grad_params (nsets x 5, may be null), and optionally the Jacobian
 * d est[p, i] / d params[p] into jac (nsets x n x 5).  The adjoint reuses the
 * forward factor and recomputes covariance derivatives. */
int hpglx_knn(const float* xyz, int n, int k, int* nbr, void* stream);
size_t hpglx_loo_state_size(int n, int k, int nsets);
int hpglx_loo_forward(const float* xyz, const float* values, int n, const int* nbr, int k, const double* params,
                      int nsets, int model, double* est, int* status, void* state, void* stream);
int hpglx_loo_backward(const float* xyz, const float* values, int n, const int* nbr, int k, const double* params,
                       int nsets, int model, const void* state, const double* grad_est, double* grad_params,
                       double* jac, void* stream);

const char* hpglx_last_error(void);
const char* hpglx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* HPGL_MI355X_HPGLX_H */
