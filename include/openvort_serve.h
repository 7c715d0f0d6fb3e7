/*
 * openvort_serve.h -- request loop over the engine of openvort_gpu.h.
 *
 * Simulation jobs (seeded ring tangles, each with its own box, point
 * spacing, temperature, counterflow and step count) arrive over time.  The
 * loop admits them in arrival order while their estimated device footprint
 * fits under an HBM cap, keeps every admitted tangle's state resident in
 * its own engine context (own HIP stream), and lets `workers` host threads
 * step the admitted jobs concurrently: each worker takes a job that is not
 * being stepped, runs one full step of the reference driver's order
 * (src/main.c:127-138: reconnect, small loops, smoothing on schedule, wrap,
 * remesh, small loops, update_tangle, rk4_step, wrap) and puts it back, so
 * jobs of different sizes overlap on the device and a finished job's slot
 * and memory go to the next arrival at once.  Each job's arithmetic is that
 * of a context of its own: its result does not depend on what else runs.
 *
 * There is no counterpart in the reference (one tangle per process).
 */
#ifndef OPENVORT_SERVE_H
#define OPENVORT_SERVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ovs_job {
    uint64_t seed;        /* ring stream (ovg_generate_rings)               */
    int32_t  n_rings;
    int32_t  steps;       /* RK4 steps requested                            */
    double   D;           /* periodic cube side, cm                         */
    double   r_min, r_max;/* ring radii, cm                                 */
    double   dxi;         /* point spacing, cm                              */
    double   alpha, alpha_p;
    double   v_ns;        /* counterflow, normal fluid along x, cm/s        */
    double   dt;          /* s                                              */
    double   arrival_s;   /* seconds after the start of the run             */
    int32_t  smooth_every;/* smoothing every k-th step, 0 = never           */
    int32_t  precision;   /* OVG_PREC_*                                     */
    int32_t  traverse;    /* OVG_TRAVERSE_*                                 */
    int32_t  leaf_size;
    int32_t  points_per_wave;
    int32_t  rec_velocity;/* OVG_RECVEL_*                                   */
    double   theta;
    int32_t  topology;    /* OVG_TOPO_* (openvort_gpu.h)                    */
    int32_t  pad_;
} ovs_job;

typedef struct ovs_result {
    int32_t  points_initial, points_final;
    int32_t  steps_done;
    int32_t  failed;       /* nonzero: an engine call failed (code)         */
    double   admit_s;      /* admission time after the start of the run     */
    double   done_s;       /* completion time                               */
    double   latency_s;    /* done - arrival (queueing included)            */
    double   line_length;  /* cm, at the end                                */
    uint64_t checksum;     /* FNV-1a over final positions, links and status */
    double   point_steps;  /* sum over steps of live points                 */
} ovs_result;

typedef struct ovs_report {
    double  wall_s;
    double  point_steps;          /* all jobs                               */
    double  point_steps_per_s;
    double  step_p50_ms, step_p99_ms;   /* one job's step, wall time         */
    double  job_p50_s, job_p99_s;       /* arrival to completion             */
    double  peak_bytes;           /* largest admitted footprint estimate     */
    int32_t jobs_done, jobs_failed;
    int32_t peak_active;
    int32_t steps_total;
    int32_t pad_;
} ovs_report;

/* Estimated device bytes of a job of n_slots slots (state, tree, scratch). */
double ovs_job_bytes(int n_slots);

/*
 * Run the jobs (sorted by arrival_s) to completion.  realtime != 0 honours
 * the arrival times on the wall clock; 0 makes every job available at once
 * (throughput under saturation).  hbm_cap_bytes bounds the sum of the
 * admitted jobs' estimates; max_active bounds their number.  results may
 * be NULL.  Returns OVG_OK, or an error code when the arguments are bad; a
 * failing job is reported in its result and the run continues.
 */
int ovs_run(const ovs_job *jobs, int n_jobs, int workers, int max_active,
            double hbm_cap_bytes, int realtime, ovs_report *report, ovs_result *results);

#ifdef __cplusplus
}
#endif

#endif /* OPENVORT_SERVE_H */
