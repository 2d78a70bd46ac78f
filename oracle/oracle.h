/*
 * oracle.h - FP64 host oracle of the VT-Physics particle-fluid step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness reference for
 * the HIP implementation and the timed CPU baseline; it is never linked into
 * or called by the product library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg load it.
 *
 * It restates the reference kernels in double precision, keeping the
 * reference's own data structure: particle arrays in attach order, a sorted
 * index array and neighbour lists neighbors[p*M + t] with self first.  The
 * discrete decisions (cell of a particle, neighbour test) are made on the
 * float-rounded positions with the same FP32 arithmetic as the reference.
 *
 * Parity status of each part is stated in the header of oracle.cpp.
 */
#ifndef VTP_ORACLE_H
#define VTP_ORACLE_H

#include <stdint.h>

#include "../include/vtp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* Objects = attach ranges [start, end) with a material, as attachObject
 * records them (PBFSolver.cpp:193-218). */
orc_sim *orc_create(const vtp_params *p, uint32_t n, const float *pos, const float *vel,
                    const float *vol_frac, const uint32_t *scene, uint32_t nobj,
                    const uint32_t *obj_start, const uint32_t *obj_end, const int32_t *obj_mat);
int orc_initialize(orc_sim *s);
int orc_step(orc_sim *s, uint32_t nsteps);
/* Field ids are vtp_field values; output in attach order as doubles, with
 * the same element counts as vtp_read. */
int orc_read(orc_sim *s, int32_t field, double *out);
/* POS, VEL (any time), VEL_PHASE [n][P][3] and Q [n][9] (after initialize). */
int orc_write(orc_sim *s, int32_t field, const double *src);
/* Test probes: orc_read of the IMM-CT rheological factors beta, kappa and
 * the polymer blocking sum of the last tick (one double per particle); and
 * orc_set_stop(1): a DFSPH tick returns right after its divergence solve. */
enum { ORC_PROBE_BETA = 100, ORC_PROBE_KAPPA_CT = 101, ORC_PROBE_BLOCKING = 102 };
void orc_set_stop(orc_sim *s, int stage);
/* Reference-layout UGNS state of the last search (sizes N*M, N, N, C, C). */
int orc_neighbors(orc_sim *s, uint32_t *lists, uint32_t *num, uint32_t *sorted, uint32_t *cell_start,
                  uint32_t *cell_end);
/* Neighbour search of the current positions only (no step). */
int orc_search(orc_sim *s);
/* iterations of the last step: div, incomp, phase; residuals at exit. */
int orc_stats(orc_sim *s, uint32_t *div_it, uint32_t *inc_it, uint32_t *phase_it, double *div_res,
              double *inc_res, uint32_t *crash);
void orc_destroy(orc_sim *s);
/* Kernel probes: W(r), gradW(r), Akinci C(|r|) for smoothing radius h. */
void orc_kernel(double h, const double *r, double *w, double *grad, double *tension_c);
int orc_cell_coord(float x, float lb, float cell);
int orc_neighbor_test(const float *a, const float *b, float cell);
/* OpenMP threads used by the oracle (0 = runtime default). */
void orc_set_threads(int n);
int orc_get_threads(void);

/* PBF adjoint in FP64 (same split as the C ABI): a taped forward of nsteps
 * ticks from the current state, the density loss at the current state, and
 * the reverse pass.  grad_x/grad_v: dL/d(final x, v) in attach order (3 per
 * particle, may be NULL); outputs dL/dv0, dL/dx0 (may be NULL) and dL/dk. */
int orc_pbf_rollout_fwd(orc_sim *s, uint32_t nsteps);
int orc_pbf_rollout_bwd(orc_sim *s, const double *grad_x, const double *grad_v, double *grad_v0,
                        double *grad_x0, double *grad_k);
/* Mean squared density deviation over the last neighbour lists and dL/dx. */
int orc_pbf_density_loss(orc_sim *s, double *loss, double *grad_x);
int orc_pbf_set_xsph(orc_sim *s, double k);

#ifdef __cplusplus
}
#endif
#endif
