/*
 * ovoracle.h -- CPU oracle for the vortex-filament hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference algorithm (robertjurco/OpenVort, src/tangle.c, src/octree.c,
 * src/boundaries.c, src/vortex_dynamics.c, src/external_velocity.c), used
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  Nothing in the product links, loads or calls it.
 *
 * Data layout is the reference's slot layout (see include/openvort_gpu.h):
 * vnodes double[3n], connections int[2n] (forward, reverse),
 * status int[2n] (status, pin_wall).
 *
 * Pinning of each function (closed forms, invariants, brute force and the
 * compiled reference) is listed in tests/test_oracle.py.  Parity unpinned:
 * none of the functions below; the smoothing restatement lives in
 * oracle/smooth.py and is pinned there.
 */
#ifndef OVORACLE_H
#define OVORACLE_H

#include "../include/openvort_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 5-point tangents and normals for every non-EMPTY slot. */
void ovo_tangents_normals(int n, const double *vnodes, const int *conn,
                          const int *status, const ovg_params *p,
                          double *tangents, double *normals);

/* Velocity of every non-EMPTY slot (direct sum or reference Barnes-Hut by
 * p->method; BH uses the reference's split-until-one tree).  tangents and
 * normals are inputs.  interactions (may be NULL) receives the number of
 * pair-formula evaluations. */
void ovo_velocities(int n, const double *vnodes, const int *conn,
                    const int *status, const ovg_params *p, double t,
                    const double *tangents, const double *normals,
                    double *vs, double *vels, int nthreads,
                    double *interactions);

/* Velocity of one slot with the direct sum (update_velocity with
 * tree == NULL). */
void ovo_velocity_direct_one(int n, const double *vnodes, const int *conn,
                             const int *status, const ovg_params *p, double t,
                             const double *tangents, const double *normals,
                             int k, double vs[3], double vels[3]);

/* Biot-Savart part only (no LIA, flows or friction) at one slot, direct. */
void ovo_bs_direct_one(int n, const double *vnodes, const int *conn,
                       const ovg_params *p, int k, double out[3]);

/* Single pair: segment [r1, r2] (r2 already pwrapped) at point r,
 * reference guards included (src/tangle.c:525-560). */
void ovo_segment_field(const double r1[3], const double r2[3],
                       const double r[3], double kappa, double out[3]);

/* One RK4 step in place on vnodes; vels holds k1 (the velocities of the
 * preceding update_tangle) and is not modified. */
void ovo_rk4_step(int n, double *vnodes, const int *conn, const int *status,
                  const ovg_params *p, double t, double dt,
                  const double *vels, int nthreads);

/* enforce_periodic_boundaries */
void ovo_enforce_periodic(int n, double *vnodes, const int *status,
                          const ovg_params *p);

/* Candidate pairs (k<l) of the reconnection scan before any reconnection,
 * brute force over all pairs.  Writes up to max_pairs pairs, returns the
 * total count. */
long ovo_reconnect_candidates(int n, const double *vnodes, const int *conn,
                              const int *status, const ovg_params *p,
                              long max_pairs, int *pairs);

/* Full reconnect pass (src/vortex_dynamics.c:186-288); conn is updated in
 * place.  Returns the number of reconnections. */
int ovo_reconnect(int n, const double *vnodes, int *conn, const int *status,
                  const ovg_params *p, double t, int nthreads);

/* Sum of |seg_pwrap(x_k, x_fwd(k))| over non-EMPTY slots with fwd >= 0. */
double ovo_line_length(int n, const double *vnodes, const int *conn,
                       const int *status, const ovg_params *p);

/* Reference-tree statistics: node count (including the empty children the
 * reference allocates) and depth. */
void ovo_bh_tree_stats(int n, const double *vnodes, const int *conn,
                       const int *status, const ovg_params *p,
                       long *nodes, int *depth, long *nonempty);

/* eliminate_small_loops (closed loops); conn/status in place; returns the
 * number of deleted points. */
int ovo_eliminate_small_loops(int n, int *conn, int *status, int cutoff);

/* remesh (src/tangle.c:1018-1061) with add_point/remove_point and the
 * reference's slot allocation (next_empty hint, capacity doubling of *N up
 * to cap).  Returns points added (-1 on overflow of cap). */
int ovo_remesh(int *N, int cap, double *vnodes, int *conn, int *status,
               int *next_empty, const ovg_params *p, double min_dist,
               double max_dist, int *removed);

/* octree_get_vs (reference tree) at m arbitrary points. */
void ovo_bh_field(int n, const double *vnodes, const int *conn,
                  const int *status, const ovg_params *p, int m,
                  const double *pts, double *out);

#ifdef __cplusplus
}
#endif
#endif
