/*
 * ref_harness.c -- drives the compiled reference (robertjurco/OpenVort
 * src/*.c, built in place by oracle/Makefile into oracle/_ref/) from
 * Python through ctypes.  TEST INFRASTRUCTURE ONLY: used by tests/ to pin
 * the oracle and by bench.py's cpu_baseline leg.
 *
 * It sets the reference globals (include/vortex_constants.h:11-60,
 * include/external_velocity.h:29-33), builds a struct tangle_state from
 * lossless binary positions, runs the reference hot-path functions and
 * copies the results back.  Linked with -Wl,--wrap=malloc so that every
 * reference malloc returns zeroed memory: octree_init leaves children[]
 * uninitialised (src/octree.c:26-35) and octree_update_means /
 * octree_destroy walk it (src/octree.c:46-49,285).
 */
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <time.h>
#include <math.h>

#include "tangle.h"
#include "utils.h"
#include "vortex_constants.h"
#include "vortex_dynamics.h"
#include "external_velocity.h"
#include "vortex_injection.h"
#include "octree.h"
#include "../include/openvort_gpu.h"

extern double rec_dist;
extern double reconnection_angle_cutoff;

#ifndef OVREF_NO_WRAP
void *__real_malloc(size_t n);
void *__wrap_malloc(size_t n) { return calloc(1, n ? n : 1); }
#endif

static void set_flow(struct ext_vel_param *dst, const ovg_flow *src)
{
    dst->type = src->type;
    dst->strength = src->strength;
    dst->time_of_linear_decrease = src->time_of_linear_decrease;
    dst->starting_strength = src->starting_strength;
    for (int k = 0; k < 3; ++k) dst->direction.p[k] = src->direction[k];
}

/* Set the reference globals from p.  Returns 0, or -1 when p names an
 * image table the reference does not have. */
int ref_setup(const ovg_params *p, int nthreads)
{
    KAPPA = p->kappa;
    VORTEX_WIDTH = p->vortex_width;
    use_mutual_friction = p->use_mutual_friction;
    alpha = p->alpha;
    alpha_p = p->alpha_p;
    use_BH = p->method == OVG_METHOD_BH;
    BH_resolution = p->theta;
    rec_dist = p->rec_dist;
    reconnection_angle_cutoff = p->rec_angle;
    freq_to_cutoff = p->freq_to_cutoff;
    global_num_threads = nthreads > 0 ? nthreads : 1;
    set_flow(&vn_conf, &p->vn);
    set_flow(&vs_conf, &p->vs);
    set_flow(&vb_conf, &p->vb);
    return 0;
}

static int set_box(struct tangle_state *t, const ovg_params *p)
{
    for (int k = 0; k < 3; ++k) {
        t->box.bottom_left_back.p[k] = p->box_lo[k];
        t->box.top_right_front.p[k] = p->box_hi[k];
    }
    for (int k = 0; k < 6; ++k) t->box.wall[k] = (wall_type)p->wall[k];
    for (int j = 0; j < p->n_images && j < 26; ++j)
        if (p->image_reflect[j]) {
            t->bimg = open_boundaries;  /* mirror tables: ref_set_wall_images */
            return 0;
        }
    switch (p->n_images) {
    case 0: t->bimg = open_boundaries; break;
    case 6: t->bimg = periodic_6; break;
    case 18: t->bimg = periodic_18; break;
    case 26: t->bimg = periodic_26; break;
    default: return -1;
    }
    return 0;
}

/* Copy of the reference's periodic image table with n entries, for the
 * test that the engine's table is the same. */
int ref_image_table(int n, int *out)
{
    const struct boundary_images *b = n == 6 ? &periodic_6 : n == 18 ? &periodic_18
                                    : n == 26 ? &periodic_26 : NULL;
    if (!b) return -1;
    for (int j = 0; j < b->n; ++j)
        for (int k = 0; k < 3; ++k) out[3 * j + k] = b->images[j].shift[k];
    return b->n;
}

/* The reference's mirror-wall image table for an OVG_IMAGES_WALL_* id
 * (src/boundary_images.c:255-302); NULL for an unknown id. */
static const struct boundary_images *wall_table(int config)
{
    switch (config) {
    case OVG_IMAGES_WALL_1_OPEN: return &wall_1_open;
    case OVG_IMAGES_WALL_1_6: return &wall_1_6;
    case OVG_IMAGES_WALL_1_18: return &wall_1_18;
    case OVG_IMAGES_WALL_1_26: return &wall_1_26;
    case OVG_IMAGES_WALL_2_4: return &wall_2_4;
    case OVG_IMAGES_WALL_2_2: return &wall_2_2;
    case OVG_IMAGES_WALL_PARALLEL_PLANES: return &wall_parallel_planes;
    case OVG_IMAGES_WALL_CANAL: return &wall_canal;
    default: return NULL;
    }
}

/* Entries actually stored in a wall table: wall_1_6 declares n = 6 over a
 * 5-entry array (reading the sixth is out of bounds), so only 5 are read. */
static int wall_table_len(int config, const struct boundary_images *b)
{
    return config == OVG_IMAGES_WALL_1_6 && b->n > 5 ? 5 : b->n;
}

/* Copy of a reference wall table: shifts and the face mask of
 * reflect[0 .. number_of_reflects).  Returns the entry count or -1. */
int ref_wall_table(int config, int *shift, int *mask)
{
    const struct boundary_images *b = wall_table(config);
    if (!b) return -1;
    int n = wall_table_len(config, b);
    for (int j = 0; j < n; ++j) {
        mask[j] = 0;
        for (int k = 0; k < 3; ++k) shift[3 * j + k] = b->images[j].shift[k];
        for (int r = 0; r < b->images[j].number_of_reflects; ++r)
            mask[j] |= 1 << b->images[j].reflect[r];
    }
    return n;
}

/* Use a reference wall table for tangle h (its walls come from the params
 * given to ref_tangle_new).  wall_1_6 is cut to its 5 stored entries. */
int ref_set_wall_images(void *h, int config)
{
    struct tangle_state *t = (struct tangle_state *)h;
    const struct boundary_images *b = wall_table(config);
    if (!b) return -1;
    t->bimg = *b;
    t->bimg.n = wall_table_len(config, b);
    return 0;
}

void *ref_tangle_new(int n, const double *vnodes, const int *conn,
                     const int *status, const ovg_params *p)
{
    struct tangle_state *t = (struct tangle_state *)calloc(1, sizeof(*t));
    create_tangle(t, (size_t)n);
    if (set_box(t, p)) { free_tangle(t); free(t); return NULL; }
    int empty = 0;
    for (int k = 0; k < n; ++k) {
        for (int c = 0; c < 3; ++c) t->vnodes[k].p[c] = vnodes[3 * k + c];
        t->connections[k].forward = conn[2 * k];
        t->connections[k].reverse = conn[2 * k + 1];
        t->status[k].status = (node_status_t)status[2 * k];
        t->status[k].pin_wall = (boundary_faces)status[2 * k + 1];
        t->recalculate[k] = 0;
        for (int c = 0; c < 3; ++c) {
            t->vels[k].p[c] = 0; t->vs[k].p[c] = 0;
            t->tangents[k].p[c] = 0; t->normals[k].p[c] = 0;
        }
        empty += status[2 * k] == EMPTY;
    }
    t->total_empty = empty;
    t->next_empty = -1;
    return t;
}

void ref_tangle_free(void *h)
{
    struct tangle_state *t = (struct tangle_state *)h;
    free_tangle(t);
    free(t);
}

void ref_set_vels(void *h, const double *vels)
{
    struct tangle_state *t = (struct tangle_state *)h;
    for (int k = 0; k < t->N; ++k)
        for (int c = 0; c < 3; ++c) t->vels[k].p[c] = vels[3 * k + c];
}

void ref_set_tangents(void *h, const double *tan, const double *nor)
{
    struct tangle_state *t = (struct tangle_state *)h;
    for (int k = 0; k < t->N; ++k)
        for (int c = 0; c < 3; ++c) {
            t->tangents[k].p[c] = tan[3 * k + c];
            t->normals[k].p[c] = nor[3 * k + c];
        }
}

void ref_get(void *h, double *vnodes, double *vels, double *vs, double *tan,
             double *nor, int *conn)
{
    struct tangle_state *t = (struct tangle_state *)h;
    for (int k = 0; k < t->N; ++k) {
        for (int c = 0; c < 3; ++c) {
            if (vnodes) vnodes[3 * k + c] = t->vnodes[k].p[c];
            if (vels) vels[3 * k + c] = t->vels[k].p[c];
            if (vs) vs[3 * k + c] = t->vs[k].p[c];
            if (tan) tan[3 * k + c] = t->tangents[k].p[c];
            if (nor) nor[3 * k + c] = t->normals[k].p[c];
        }
        if (conn) {
            conn[2 * k] = t->connections[k].forward;
            conn[2 * k + 1] = t->connections[k].reverse;
        }
    }
}

extern int small_loop_cutoff;

void ref_eliminate_small_loops(void *h, int cutoff)
{
    small_loop_cutoff = cutoff;
    eliminate_small_loops((struct tangle_state *)h);
}

void ref_remesh(void *h, double min_dist, double max_dist)
{
    remesh((struct tangle_state *)h, min_dist, max_dist, 0.0);
}

/* enforce_wall_boundaries (src/tangle.c:966-1016) at time t, new wall nodes
 * of status pin (the pin_mode global, src/initial_conditions.c:54) */
void ref_enforce_walls(void *h, double t, int pin)
{
    pin_mode = pin;
    enforce_wall_boundaries((struct tangle_state *)h, t);
}

int ref_num_slots(void *h) { return ((struct tangle_state *)h)->N; }

/* save_tangle / load_tangle, src/utils.c:132-199; the load goes into a
 * tangle made as src/main.c:84-86 does (create_tangle(512)) with the box
 * and image table of p. */
void ref_save_tangle(void *h, const char *path) { save_tangle(path, (struct tangle_state *)h); }

void *ref_load_tangle(const char *path, const ovg_params *p)
{
    struct tangle_state *t = (struct tangle_state *)calloc(1, sizeof(*t));
    create_tangle(t, 512);
    if (set_box(t, p)) { free_tangle(t); free(t); return NULL; }
    load_tangle(path, t);
    return t;
}

void ref_get_status(void *h, int *st)
{
    struct tangle_state *t = (struct tangle_state *)h;
    for (int k = 0; k < t->N; ++k) {
        st[2 * k] = t->status[k].status;
        st[2 * k + 1] = t->status[k].pin_wall;
    }
}

void ref_update_tangents_normals(void *h) { update_tangents_normals((struct tangle_state *)h); }
void ref_update_velocities(void *h, double t) { update_velocities((struct tangle_state *)h, t); }
void ref_update_tangle(void *h, double t) { update_tangle((struct tangle_state *)h, t); }
void ref_rk4_step(void *h, double t, double dt) { rk4_step((struct tangle_state *)h, t, dt); }
int ref_reconnect(void *h, double t) { return reconnect((struct tangle_state *)h, t); }
void ref_enforce_periodic(void *h) { enforce_periodic_boundaries((struct tangle_state *)h); }

/* update_velocity for one slot with tree == NULL (the direct sum). */
void ref_velocity_one(void *h, int k, double t, double *vs, double *vels)
{
    struct tangle_state *tg = (struct tangle_state *)h;
    update_velocity(tg, k, t, NULL);
    for (int c = 0; c < 3; ++c) {
        vs[c] = tg->vs[k].p[c];
        vels[c] = tg->vels[k].p[c];
    }
}

/* octree_get_vs at arbitrary points, one tree build. */
void ref_bh_field(void *h, int m, const double *pts, double theta, double *out)
{
    struct tangle_state *tg = (struct tangle_state *)h;
    struct octree *tree = octree_build(tg);
    for (int i = 0; i < m; ++i) {
        struct vec3 r = {{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}}, v;
        octree_get_vs(tg, tree, &r, theta, &v);
        for (int c = 0; c < 3; ++c) out[3 * i + c] = v.p[c];
    }
    octree_destroy(tree);
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Wall-clock seconds of the reference's per-step phases on this tangle:
 * [0] update_tangle, [1] rk4_step, [2] one octree_build + destroy. */
void ref_time_phases(void *h, double t, double dt, double *sec)
{
    struct tangle_state *tg = (struct tangle_state *)h;
    double t0 = now_s();
    update_tangle(tg, t);
    double t1 = now_s();
    rk4_step(tg, t, dt);
    double t2 = now_s();
    struct octree *tree = octree_build(tg);
    octree_destroy(tree);
    double t3 = now_s();
    sec[0] = t1 - t0;
    sec[1] = t2 - t1;
    sec[2] = t3 - t2;
}

/* add_loop (src/vortex_injection.c:44-103) with the driver's dl_min/dl_max
 * globals set to the given values. */
void ref_add_loop(void *h, const double *center, const double *dir, double r, double dl_min,
                  double dl_max)
{
    global_dl_min = dl_min;
    global_dl_max = dl_max;
    struct vec3 c = vec3(center[0], center[1], center[2]);
    struct vec3 d = vec3(dir[0], dir[1], dir[2]);
    add_loop((struct tangle_state *)h, &c, &d, r);
}

/* srand48(seed), then inject_loop (src/vortex_injection.c:711-745) at each
 * of the n times.  inject_loop keeps its last injection time in a static
 * that starts at 0 when the library is loaded and persists across calls. */
/* srand48(seed), then inject_line_pairs (src/vortex_injection.c:760-800) at
 * each time with global_dl_min / global_dl_max set (its static last
 * injection time persists for the life of the library) */
void ref_inject_line_pairs(void *h, long seed, const double *times, int n, double frequency,
                           int n_pairs, int polarized, double dl_min, double dl_max, int pin)
{
    struct tangle_state *t = (struct tangle_state *)h;
    global_dl_min = dl_min;
    global_dl_max = dl_max;
    pin_mode = pin;
    srand48(seed);
    for (int i = 0; i < n; ++i) inject_line_pairs(t, times[i], frequency, n_pairs, polarized);
}

void ref_inject_loops(void *h, long seed, const double *times, int n, double frequency,
                      double dl_min, double dl_max)
{
    global_dl_min = dl_min;
    global_dl_max = dl_max;
    srand48(seed);
    for (int i = 0; i < n; ++i) inject_loop((struct tangle_state *)h, times[i], frequency);
}
