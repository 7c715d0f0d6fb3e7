/*
 * dropin_driver.c -- the reference driver's step loop (src/main.c:117-141,
 * without vortex injection and wall handling) over the reference's own
 * tangle API.  TEST INFRASTRUCTURE ONLY.
 *
 * oracle/Makefile links it twice:
 *   _ref/ovref_driver   with the reference objects as they are: every call
 *                       runs the reference's CPU code;
 *   _ref/ovref_dropin   with the same objects compiled so that the hot-path
 *                       definitions (update_tangents_normals,
 *                       update_velocities, update_velocity, rk4_step,
 *                       reconnect, enforce_periodic_boundaries) are renamed
 *                       away, so the unmodified calls below resolve to
 *                       libopenvort_dropin.so (the GPU), while remesh and
 *                       eliminate_small_loops stay the reference's CPU code.
 * smooth() (src/vortex_smooth.c, FFTW) exists only in the drop-in link; the
 * CPU link runs with smoothing off.
 *
 * Usage: <driver> in.bin out.bin steps dt smooth_every dl_min dl_max
 *   in.bin:  int32 n, ovg_params, double vnodes[3n], int32 conn[2n],
 *            int32 status[2n]
 *   out.bin: int32 n, double vnodes[3n], int32 conn[2n], int32 status[2n],
 *            double vels[3n]
 */
#include <stdio.h>
#include <stdlib.h>

#include "tangle.h"
#include "vortex_dynamics.h"
#include "../include/openvort_gpu.h"

int ref_setup(const ovg_params *p, int nthreads);
void *ref_tangle_new(int n, const double *vnodes, const int *conn, const int *status,
                     const ovg_params *p);
#ifdef DRIVER_SMOOTH
void smooth(struct tangle_state *tangle);
#endif

static int rd(void *p, size_t sz, size_t n, FILE *f) { return fread(p, sz, n, f) == n; }

int main(int argc, char **argv)
{
    if (argc < 8) {
        fprintf(stderr, "usage: %s in.bin out.bin steps dt smooth_every dl_min dl_max\n",
                argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 2; }
    int n = 0;
    ovg_params p;
    if (!rd(&n, sizeof n, 1, f) || !rd(&p, sizeof p, 1, f)) return 2;
    /* the file must hold exactly this build's ovg_params layout: a writer
     * with another layout would hand the reference shifted links */
    long here = ftell(f);
    fseek(f, 0, SEEK_END);
    long want = here + (long)n * (long)(3 * sizeof(double) + 4 * sizeof(int));
    if (n < 0 || ftell(f) != want) {
        fprintf(stderr, "%s: %ld bytes, expected %ld for %d points and a %zu-byte ovg_params\n",
                argv[1], ftell(f), want, n, sizeof p);
        return 2;
    }
    fseek(f, here, SEEK_SET);
    double *x = malloc(sizeof(double) * 3 * (size_t)n);
    int *conn = malloc(sizeof(int) * 2 * (size_t)n), *st = malloc(sizeof(int) * 2 * (size_t)n);
    if (!rd(x, sizeof(double), 3 * (size_t)n, f) || !rd(conn, sizeof(int), 2 * (size_t)n, f) ||
        !rd(st, sizeof(int), 2 * (size_t)n, f))
        return 2;
    fclose(f);
    const int steps = atoi(argv[3]), smooth_every = atoi(argv[5]);
    const double dt = atof(argv[4]), dl_min = atof(argv[6]), dl_max = atof(argv[7]);
    ref_setup(&p, 8);
    struct tangle_state *t = ref_tangle_new(n, x, conn, st, &p);
    if (!t) return 3;
    double time = 0;
    for (int k = 0; k < steps; ++k) {
        eliminate_small_loops(t);
        reconnect(t, time);
        eliminate_small_loops(t);
#ifdef DRIVER_SMOOTH
        if (smooth_every > 0 && k % smooth_every == 0) smooth(t);
#else
        (void)smooth_every;
#endif
        enforce_periodic_boundaries(t);
        remesh(t, dl_min, dl_max, time);
        eliminate_small_loops(t);
        update_tangle(t, time);
        rk4_step(t, time, dt);
        enforce_periodic_boundaries(t);
        time += dt;
    }
    f = fopen(argv[2], "wb");
    if (!f) { perror(argv[2]); return 2; }
    n = t->N;
    fwrite(&n, sizeof n, 1, f);
    for (int k = 0; k < n; ++k) fwrite(t->vnodes[k].p, sizeof(double), 3, f);
    for (int k = 0; k < n; ++k) {
        int c[2] = {t->connections[k].forward, t->connections[k].reverse};
        fwrite(c, sizeof(int), 2, f);
    }
    for (int k = 0; k < n; ++k) {
        int s[2] = {(int)t->status[k].status, (int)t->status[k].pin_wall};
        fwrite(s, sizeof(int), 2, f);
    }
    for (int k = 0; k < n; ++k) fwrite(t->vels[k].p, sizeof(double), 3, f);
    fclose(f);
    printf("steps %d points %d\n", steps, n);
    return 0;
}
