/*
 * openvort_dropin.h -- drop-in replacements, with the reference's exact C
 * prototypes, for OpenVort's hot-path functions, implemented on the GPU
 * through the C ABI of openvort_gpu.h (library: libopenvort_dropin.so).
 *
 * A reference build links these instead of its own definitions, ahead
 * of the reference's objects on the link line.  They read the reference's globals
 * (include/vortex_constants.h:11-60, include/external_velocity.h:29-33),
 * which the reference's src/initial_conditions.c defines, and the
 * struct tangle_state of the caller, whose layout is restated below
 * (include/tangle.h:39-66, include/boundaries.h:40-81,
 * include/vec3_math.h:16-18).
 *
 * Each call uploads the tangle, runs on the GPU and writes the results
 * back into the caller's arrays, so the caller keeps owning its memory and
 * may reallocate it between calls (expand_tangle, src/tangle.c:68-101).
 * Errors print one line to stderr and abort(), never silently continue.
 */
#ifndef OPENVORT_DROPIN_H
#define OPENVORT_DROPIN_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef VEC3_MATH_H
struct vec3 { double p[3]; };
#endif

#ifndef TANGLE_H
struct neighbour_t { int forward; int reverse; };
#endif

#ifndef BOUNDARIES_H
typedef enum _wt { WALL_OPEN, WALL_PERIODIC, WALL_MIRROR } wall_type;
typedef enum _fc { NOT_A_FACE = -1, LEFT, RIGHT, BACK, FRONT, DOWN, UP } boundary_faces;
typedef enum _ns_e { EMPTY = -1, FREE, PINNED, PINNED_SLIP } node_status_t;
struct node_status { node_status_t status; boundary_faces pin_wall; };
struct domain_box {
    struct vec3 bottom_left_back;
    struct vec3 top_right_front;
    wall_type wall[6];
};
struct image_tangle { int shift[3]; int reflect[3]; int number_of_reflects; };
struct boundary_images { const struct image_tangle *images; int n; };
#endif

#ifndef TANGLE_H
struct tangle_state {
    struct vec3 *vnodes;
    struct vec3 *vels;
    struct vec3 *vs;
    struct vec3 *tangents;
    struct vec3 *normals;
    int *recalculate;
    struct neighbour_t *connections;
    struct node_status *status;
    struct domain_box box;
    struct boundary_images bimg;
    int N;
    int next_empty;
    int total_empty;
};
#endif

struct octree;

/* src/tangle.c:509-516 */
void update_tangents_normals(struct tangle_state *tangle);
/* src/tangle.c:733-749 */
void update_velocities(struct tangle_state *tangle, double t);
/* src/tangle.c:600-726; the GPU evaluates the direct sum (tree == NULL,
 * the reference's own external callers, src/vortex_dynamics.c:250-251)
 * and ignores a non-NULL tree. */
void update_velocity(struct tangle_state *tangle, int k, double t, struct octree *tree);
/* src/vortex_dynamics.c:50-119 */
void rk4_step(struct tangle_state *tangle, double t, double dt);
/* src/vortex_dynamics.c:186-288 */
int reconnect(struct tangle_state *tangle, double t);
/* src/vortex_smooth.c:459-462 */
void smooth(struct tangle_state *tangle);
/* src/tangle.c:793-814 */
void enforce_periodic_boundaries(struct tangle_state *tangle);

/*
 * Engine choices the reference has no global for.  Defaults: FP64, the
 * reference's split-until-one tree with per-point traversal.  Environment
 * variables OPENVORT_PRECISION=mixed, OPENVORT_TRAVERSE=group and
 * OPENVORT_LEAF=<n> set them at first use; this call overrides them.
 */
void openvort_dropin_configure(int precision, int traverse, int leaf_size);

#ifdef __cplusplus
}
#endif

#endif /* OPENVORT_DROPIN_H */
