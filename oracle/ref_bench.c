/*
 * ref_bench.c -- CPU baseline timer around the compiled reference
 * (robertjurco/OpenVort src/*.c built with the reference Makefile's flags
 * by oracle/Makefile).  TEST/BENCH INFRASTRUCTURE ONLY: bench.py runs it as
 * a child process for its cpu_baseline leg.
 *
 * Usage: ovref_bench <tangle.bin> <threads> <steps> <dt> [reconnect 0|1] [mode]
 *
 * tangle.bin: int32 n, then ovg_params (include/openvort_gpu.h), then
 * double vnodes[3n], int32 connections[2n], int32 status[2n].
 *
 * mode "steps" (default): `steps` steps of the reference per-step order of
 *   src/main.c:127-138 restricted to the hot path: [reconnect] +
 *   update_tangle + rk4_step + enforce_periodic_boundaries.
 * mode "phases": one of each phase of a step at the file's full size:
 *   octree_build + octree_destroy (src/octree.c:139-156,41-52) once on its
 *   own, then update_tangle (include/tangle.h:114-118: tangents, tree,
 *   velocities), rk4_step (src/vortex_dynamics.c:50-119: three more
 *   evaluations; skipped when steps is 0, reported as -1) and
 *   enforce_periodic_boundaries.
 * mode "reconnect": reconnect (src/vortex_dynamics.c:186-288) once, after
 *   update_tangents_normals as in the driver's order.
 * Every phase is timed with CLOCK_MONOTONIC; one JSON line is printed.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "tangle.h"
#include "octree.h"
#include "vortex_constants.h"
#include "vortex_dynamics.h"
#include "external_velocity.h"
#include "../include/openvort_gpu.h"

int ref_setup(const ovg_params *p, int nthreads);
void *ref_tangle_new(int n, const double *vnodes, const int *conn,
                     const int *status, const ovg_params *p);

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int run_steps(struct tangle_state *t, int live, int threads, int steps, double dt,
                     int do_rec)
{
    double t_rec = 0, t_upd = 0, t_rk4 = 0, t_wrap = 0, time = 0;
    int nrec = 0;
    double *step_s = (double *)malloc(sizeof(double) * (size_t)(steps > 0 ? steps : 1));
    for (int s = 0; s < steps; ++s) {
        double a = now_s();
        if (do_rec) nrec += reconnect(t, time);
        double b = now_s();
        update_tangle(t, time);
        double c = now_s();
        rk4_step(t, time, dt);
        double d = now_s();
        enforce_periodic_boundaries(t);
        double e = now_s();
        t_rec += b - a; t_upd += c - b; t_rk4 += d - c; t_wrap += e - d;
        step_s[s] = e - a;
        time += dt;
    }
    double total = t_rec + t_upd + t_rk4 + t_wrap;
    printf("{\"mode\": \"steps\", \"points\": %d, \"threads\": %d, \"steps\": %d, "
           "\"seconds\": %.6f, \"reconnect_s\": %.6f, \"update_tangle_s\": %.6f, "
           "\"rk4_step_s\": %.6f, \"wrap_s\": %.6f, \"reconnections\": %d, "
           "\"point_steps_per_s\": %.6g, \"step_s\": [",
           live, threads, steps, total, t_rec, t_upd, t_rk4, t_wrap, nrec,
           total > 0 ? (double)live * steps / total : 0.0);
    for (int s = 0; s < steps; ++s) printf("%s%.6f", s ? ", " : "", step_s[s]);
    printf("]}\n");
    free(step_s);
    return 0;
}

static int run_phases(struct tangle_state *t, int live, int threads, double dt, int do_rk4)
{
    double a = now_s();
    struct octree *tree = octree_build(t);
    double b = now_s();
    octree_destroy(tree);  /* frees the root too (src/octree.c:41-52) */
    double c = now_s();
    update_tangle(t, 0.0);
    double d = now_s();
    if (do_rk4) rk4_step(t, 0.0, dt);
    double e = now_s();
    enforce_periodic_boundaries(t);
    double f = now_s();
    printf("{\"mode\": \"phases\", \"points\": %d, \"threads\": %d, "
           "\"tree_build_s\": %.6f, \"tree_destroy_s\": %.6f, \"update_tangle_s\": %.6f, "
           "\"rk4_step_s\": %.6f, \"wrap_s\": %.6f}\n",
           live, threads, b - a, c - b, d - c, do_rk4 ? e - d : -1.0, f - e);
    return 0;
}

static int run_reconnect(struct tangle_state *t, int live, int threads)
{
    update_tangents_normals(t);
    double a = now_s();
    int nrec = reconnect(t, 0.0);
    double b = now_s();
    printf("{\"mode\": \"reconnect\", \"points\": %d, \"threads\": %d, "
           "\"reconnect_s\": %.6f, \"reconnections\": %d}\n",
           live, threads, b - a, nrec);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s tangle.bin threads steps dt [reconnect] [mode]\n",
                argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 2; }
    int n;
    ovg_params p;
    if (fread(&n, 4, 1, f) != 1 || fread(&p, sizeof p, 1, f) != 1) return 2;
    double *x = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    int *conn = (int *)malloc(sizeof(int) * 2 * (size_t)n);
    int *st = (int *)malloc(sizeof(int) * 2 * (size_t)n);
    if (fread(x, 8, 3 * (size_t)n, f) != 3 * (size_t)n ||
        fread(conn, 4, 2 * (size_t)n, f) != 2 * (size_t)n ||
        fread(st, 4, 2 * (size_t)n, f) != 2 * (size_t)n) {
        fprintf(stderr, "short read\n");
        return 2;
    }
    fclose(f);
    int threads = atoi(argv[2]);
    int steps = atoi(argv[3]);
    double dt = atof(argv[4]);
    int do_rec = argc > 5 ? atoi(argv[5]) : 0;
    const char *mode = argc > 6 ? argv[6] : "steps";
    ref_setup(&p, threads);
    struct tangle_state *t = (struct tangle_state *)ref_tangle_new(n, x, conn, st, &p);
    if (!t) { fprintf(stderr, "bad image table\n"); return 2; }
    int live = 0;
    for (int k = 0; k < n; ++k) live += st[2 * k] != -1;
    int r;
    if (!strcmp(mode, "phases"))
        r = run_phases(t, live, threads, dt, steps > 0);
    else if (!strcmp(mode, "reconnect"))
        r = run_reconnect(t, live, threads);
    else if (!strcmp(mode, "steps"))
        r = run_steps(t, live, threads, steps, dt, do_rec);
    else {
        fprintf(stderr, "unknown mode %s\n", mode);
        r = 2;
    }
    free_tangle(t);
    free(t);
    free(x); free(conn); free(st);
    return r;
}
