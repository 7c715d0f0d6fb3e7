/*
 * openvort_gpu.h -- C ABI of the MI355X vortex-filament engine.
 *
 * The engine evaluates the superfluid velocity of a tangle of quantised
 * vortex filaments (Biot-Savart over straight segments, periodic image
 * shells, local-induction term, mutual friction) and advances it with the
 * reference's RK4 scheme, all on one GPU.  Tangle state uses the
 * reference's slot layout so that the drop-in shims in openvort_dropin.h
 * can pass struct tangle_state arrays straight through:
 *
 *   vnodes       double[3*n]  positions, x,y,z per slot
 *                             (reference: struct vec3, include/vec3_math.h:16-18)
 *   connections  int[2*n]     forward, reverse per slot, -1 = none
 *                             (reference: struct neighbour_t, include/tangle.h:39-42)
 *   status       int[2*n]     status, pin_wall per slot; status -1 = EMPTY,
 *                             0 = FREE, 1 = PINNED, 2 = PINNED_SLIP
 *                             (reference: struct node_status, include/boundaries.h:50-53)
 *
 * Every entry point returns OVG_OK (0) or a nonzero OVG_E* code; the text
 * of the last error on the calling thread is returned by ovg_last_error().
 * Nothing here prints, exits or falls back to the CPU: when no GPU or no
 * gfx950 code object is available, ovg_create fails with OVG_EDEVICE.
 *
 * Threading: one host thread drives one context at a time.  Contexts are
 * independent; each owns one HIP stream (default priority) and its device
 * buffers.
 */
#ifndef OPENVORT_GPU_H
#define OPENVORT_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define OVG_OK           0
#define OVG_EINVAL       1   /* bad argument, shape or state */
#define OVG_ENOMEM       2   /* device or host allocation failed / cap exceeded */
#define OVG_EDEVICE      3   /* HIP runtime error, no device, missing code object */
#define OVG_EUNSUPPORTED 4   /* feature outside the engine's scope (e.g. topology passes on a forest) */
#define OVG_ELIMIT       5   /* an internal fixed-size limit was exceeded */

/* ---- node status and wall kinds (reference enums) -------------------- */
#define OVG_EMPTY       (-1)  /* include/boundaries.h:50-53 */
#define OVG_FREE         0
#define OVG_PINNED       1
#define OVG_PINNED_SLIP  2

#define OVG_WALL_OPEN      0  /* include/boundaries.h:18-22 */
#define OVG_WALL_PERIODIC  1
#define OVG_WALL_MIRROR    2

/* boundary faces, include/boundaries.h:25-35 (bit f of image_reflect) */
#define OVG_FACE_LEFT   0   /* -x */
#define OVG_FACE_RIGHT  1   /* +x */
#define OVG_FACE_BACK   2   /* -y */
#define OVG_FACE_FRONT  3   /* +y */
#define OVG_FACE_DOWN   4   /* -z */
#define OVG_FACE_UP     5   /* +z */

/* image tables of src/boundary_images.c:75-302, for
 * ovg_params_set_wall_images */
#define OVG_IMAGES_WALL_1_OPEN          1  /* wall_1_open: DOWN mirror only  */
#define OVG_IMAGES_WALL_1_6             2  /* wall_1_6                       */
#define OVG_IMAGES_WALL_1_18            3  /* wall_1_18                      */
#define OVG_IMAGES_WALL_1_26            4  /* wall_1_26                      */
#define OVG_IMAGES_WALL_2_4             5  /* wall_2_4: z walls, x/y periodic */
#define OVG_IMAGES_WALL_2_2             6  /* wall_2_2: y, z walls           */
#define OVG_IMAGES_WALL_PARALLEL_PLANES 7  /* wall_parallel_planes           */
#define OVG_IMAGES_WALL_CANAL           8  /* wall_canal: x periodic         */

/* external velocity kinds, include/external_velocity.h:27-31 */
#define OVG_FLOW_NONE        0
#define OVG_FLOW_SIMPLE      1
#define OVG_FLOW_DECREASING  2

/* Biot-Savart evaluation method */
#define OVG_METHOD_DIRECT  0  /* all segments, every image (src/tangle.c:652-687) */
#define OVG_METHOD_BH      1  /* Barnes-Hut octree (src/octree.c:296-356)        */

/* arithmetic of the pair interactions */
#define OVG_PREC_FP64   0     /* every pair in FP64, reference operation order */
#define OVG_PREC_MIXED  1     /* far-field pairs in FP32, per-point sums in FP64 */

/* Barnes-Hut traversal organisation */
#define OVG_TRAVERSE_POINT  0 /* one lane per (point, image), the reference's
                                 per-point recursion made stackless             */
#define OVG_TRAVERSE_GROUP  1 /* one wavefront per 64 Morton-adjacent points and
                                 one image; same per-point acceptance test       */
#define OVG_TRAVERSE_LIST   2 /* one wavefront per points_per_wave Morton-adjacent
                                 points: one tree walk per group and image with
                                 the acceptance test applied to the group's
                                 bounding box (never coarser than the per-point
                                 test), into an interaction list in LDS that
                                 every lane then evaluates                       */

/* velocity source of the reconnection approach test */
#define OVG_RECVEL_DIRECT  0
#define OVG_RECVEL_TREE    1

/* Where the topology passes run (ovg_params.topology).
 * OVG_TOPO_HOST: the sequential passes walk host mirrors of the links and
 *   send the changed slots back (every tangle, walls and pinned lines
 *   included).
 * OVG_TOPO_DEVICE: reconnection (candidates, endpoint velocities and the
 *   greedy pass), small-loop removal and remesh run as device kernels on
 *   the links in HBM and queue without waiting for the host; results are
 *   identical to OVG_TOPO_HOST.  Taken by one periodic tangle without
 *   pinned nodes or mirror images whose reconnection velocities are the
 *   direct sum or the list traversal; any other tangle runs the host
 *   passes.  The point count and the counts the passes return are read
 *   back when a later call needs them (one wait per step in the reference
 *   driver's order), or at once when a count pointer is given; with
 *   direct-sum reconnection velocities the pass also waits once for its
 *   pair and endpoint counts (to plan the direct sum), and a remesh on a
 *   tangle with few free slots waits once for a bound on its insertions
 *   (the slots double only when that may not fit).  A device
 *   pass that meets more than its capacity of candidate pairs (grown
 *   automatically from the counts seen, up to 32768 pairs and 65536
 *   endpoints) or finds no empty slot for an insertion stops and the next
 *   call that waits returns OVG_ELIMIT. */
#define OVG_TOPO_HOST    0
#define OVG_TOPO_DEVICE  1

/* External (prescribed) flow, include/external_velocity.h:11-19. */
typedef struct ovg_flow {
    int32_t type;                 /* OVG_FLOW_*                               */
    int32_t pad_;
    double  strength;             /* cm/s                                     */
    double  time_of_linear_decrease;
    double  starting_strength;
    double  direction[3];         /* normalised on every use, as the reference */
} ovg_flow;

/*
 * Physics, geometry and method of one tangle.  Field names follow the
 * reference globals of include/vortex_constants.h:11-60.
 */
typedef struct ovg_params {
    double  kappa;                /* KAPPA, cm^2/s                            */
    double  vortex_width;         /* VORTEX_WIDTH (a0), cm                    */
    int32_t use_mutual_friction;
    int32_t pad0_;
    double  alpha, alpha_p;       /* mutual friction coefficients             */
    ovg_flow vn, vs, vb;          /* normal, superfluid, boundary flows       */

    double  box_lo[3];            /* domain_box.bottom_left_back              */
    double  box_hi[3];            /* domain_box.top_right_front               */
    int32_t wall[6];              /* LEFT RIGHT BACK FRONT DOWN UP, OVG_WALL_* */
    int32_t n_images;             /* entries of images[] / image_reflect[]    */
    int32_t images[26][3];        /* integer box shifts, reference order      */

    int32_t method;               /* OVG_METHOD_*                             */
    int32_t precision;            /* OVG_PREC_*                               */
    double  theta;                /* BH_resolution (opening angle)            */
    int32_t leaf_size;            /* points per Barnes-Hut bucket, >= 1; 1 is the
                                     reference's split-until-one tree           */
    int32_t traverse;             /* OVG_TRAVERSE_*                           */

    double  rec_dist;             /* reconnection distance, cm                */
    double  rec_angle;            /* reconnection_angle_cutoff, radians       */
    double  freq_to_cutoff;       /* smoothing: fraction of modes removed     */
    int32_t rk4_classic;          /* 0: reference stage 4 at x+dt/2*k3
                                     (src/vortex_dynamics.c:90); 1: x+dt*k3    */
    int32_t rec_velocity;         /* velocities of the reconnection approach
                                     test: OVG_RECVEL_DIRECT (the reference's
                                     update_velocity with tree == NULL,
                                     src/vortex_dynamics.c:250-251) or
                                     OVG_RECVEL_TREE (this context's
                                     Barnes-Hut evaluator)                     */
    double  mixed_near;           /* OVG_PREC_MIXED: segment (leaf) pairs closer
                                     than this (cm) stay FP64; 0 = none        */
    int32_t points_per_wave;      /* OVG_TRAVERSE_LIST group size: 8, 16, 21,
                                     32 or 64 (0 = 16)                          */
    int32_t group_level;          /* OVG_TRAVERSE_LIST: groups never span two
                                     octree cells of this level (1..21); 0 =
                                     the coarsest level whose cells are at most
                                     8 * mixed_near wide (5 if mixed_near = 0)  */
    int32_t reuse_lists;          /* OVG_TRAVERSE_LIST: 1 = the interaction
                                     lists of ovg_update_tangle's walk serve
                                     the stages of the ovg_rk4_step that
                                     follows it (same tree, refreshed at each
                                     stage's positions; acceptance decided at
                                     the step's first stage); 0 = every stage
                                     walks its own tree (the reference's
                                     build_tree per velocity evaluation)     */
    int32_t image_reflect[26];    /* per image, the faces it mirrors through
                                     (bit OVG_FACE_*; image_tangle.reflect of
                                     include/boundaries.h:66-70): the target
                                     moves to 2 * wall - r on that axis
                                     (src/tangle.c:380-410) and, in Barnes-Hut
                                     mode, the image's velocity is reflected
                                     (src/boundaries.c:201-211); the direct
                                     mode adds it unreflected, as the
                                     reference does (src/tangle.c:678-683).
                                     0 = pure shift.  At most one face per
                                     axis.  reuse_lists takes shift images
                                     only.                                      */
    /* OVG_TRAVERSE_LIST launch shape (no counterpart in the reference; the
     * autotuner sets them).  Results differ between shapes only in the
     * FP32 summation order, except list_blocks, which changes nothing. */
    int32_t list_stack;           /* traversal stack depth in ranges, 128..384
                                     in steps of 64 (0 = 384)                   */
    int32_t list_tile;            /* far-field flush tile in entries: 96 or
                                     128 (0 = 128)                              */
    int32_t list_blocks;          /* 0: one wavefront per target group; n > 0:
                                     a persistent grid of n wavefronts per CU
                                     (1..64) striding over the groups           */
    int32_t pin_mode;             /* pin_mode (src/initial_conditions.c:54): the
                                     status of the nodes wall clipping creates,
                                     OVG_PINNED or OVG_PINNED_SLIP (0 = PINNED) */
    int32_t topology;             /* where ovg_reconnect, ovg_eliminate_small_loops
                                     and ovg_remesh run: OVG_TOPO_HOST (0) or
                                     OVG_TOPO_DEVICE (see below)               */
} ovg_params;

/* Fill p with the reference defaults (src/initial_conditions.c:12-91) for a
 * periodic cube [0,D]^3 with the 26 periodic images of
 * src/boundary_images.c:45-72.  vn/vs/vb are set to no flow. */
void ovg_params_default(ovg_params *p, double D);

/* Load the periodic image table with n = 6, 18 or 26 entries
 * (src/boundary_images.c:13-72); n = 0 clears it. */
int ovg_params_set_periodic_images(ovg_params *p, int n);

/* Load one of the reference's mirror-wall image tables (OVG_IMAGES_WALL_*,
 * src/boundary_images.c:75-302) into images / image_reflect / n_images, in
 * the reference's order and with its quirks (wall_1_18's {1,-1,0} entry
 * mirrors through LEFT; wall_parallel_planes repeats two entries).
 * wall_1_6 declares 6 entries but holds 5 (reading the sixth is out of
 * bounds in the reference): this table has the 5.  The walls themselves
 * (p->wall) are the caller's, as in the reference's set-up
 * (src/initial_conditions.c:135-165). */
int ovg_params_set_wall_images(ovg_params *p, int config);

const char *ovg_last_error(void);
const char *ovg_version(void);

/* Cap on the device memory, in bytes, that the buffers of every context of
 * this process may hold together; 0 (the default) is no cap.  A buffer that
 * would exceed it is not allocated and the call that needed it fails with
 * OVG_ENOMEM, as it would on a failed hipMalloc.  Returns the previous cap.
 * (No counterpart in the reference: a bound for processes that share the
 * GPU's memory, and a way to exercise the out-of-memory paths.) */
size_t ovg_set_device_memory_limit(size_t bytes);
/* Device bytes held by the buffers of every context of this process. */
size_t ovg_device_memory_in_use(void);

/* sizeof(ovg_params) and sizeof(ovg_stats) as compiled into the library,
 * for bindings that restate the structs (ctypes, FFI). */
size_t ovg_params_sizeof(void);
size_t ovg_stats_sizeof(void);

/* ---- context --------------------------------------------------------- */
typedef struct ovg_ctx ovg_ctx;

/* Create a context on HIP device `device`. */
int ovg_create(ovg_ctx **out, int device);
int ovg_destroy(ovg_ctx *ctx);

int ovg_set_params(ovg_ctx *ctx, const ovg_params *p);
int ovg_get_params(const ovg_ctx *ctx, ovg_params *p);

/*
 * Upload a tangle of n slots.  Replaces positions, connections and status;
 * velocities, tangents and normals become zero.  Re-upload whenever the
 * host changed capacity or topology (the reference's expand_tangle
 * reallocates, src/tangle.c:68-101).  FREE slots need both links; PINNED
 * and PINNED_SLIP slots (pin_wall 0..5, OVG_FACE_*) may end a line
 * (link -1).  A tangle with pinned slots is evaluated, stepped, clipped
 * to its walls (ovg_enforce_walls), reconnected, remeshed and smoothed as
 * the reference does (open lines: src/vortex_smooth.c:316-444).
 */
int ovg_upload(ovg_ctx *ctx, int n, const double *vnodes,
               const int *connections, const int *status);

/*
 * A forest: n_tangles independent tangles in one context, evaluated and
 * stepped together (no counterpart in the reference, which holds one
 * tangle per process; a batch of update_tangle / rk4_step calls of
 * src/tangle.c:733-749 and src/vortex_dynamics.c:50-119).  Tangle j takes
 * slots slot_offsets[j] .. slot_offsets[j+1]-1 of vnodes / connections /
 * status (each tangle in the reference layout with its OWN slot numbering:
 * its links are 0-based within its range), and params[j] (its box, images,
 * flows, friction, theta, mixed_near ...).  Every tangle takes the
 * Barnes-Hut list traversal, and precision, leaf_size, points_per_wave,
 * group_level and rk4_classic must be tangle 0's (one launch runs them
 * all); flows must not depend on time.  One tree build and one traversal
 * launch per evaluation cover every tangle (tangle index above the Morton
 * key, groups that never cross tangles, per-tangle parameter table), and
 * each tangle's results are bit-identical to the same tangle uploaded
 * alone.  ovg_update_tangle, ovg_update_velocities, ovg_update_tangents_normals,
 * ovg_rk4_step, ovg_enforce_periodic, ovg_download*, ovg_upload_positions,
 * ovg_upload_field, ovg_autotune and the graph calls take a forest (global
 * slot indices in downloads); the topology passes, ovg_velocity_at*,
 * ovg_set_params and ovg_save_tangle return OVG_EUNSUPPORTED.  ovg_upload
 * makes the context a single tangle again.
 */
#define OVG_FOREST_MAX_TANGLES 32767
int ovg_upload_forest(ovg_ctx *ctx, int n_tangles, const ovg_params *params,
                      const int *slot_offsets, const double *vnodes,
                      const int *connections, const int *status);
/* Tangles in the context (1 unless a forest was uploaded). */
int ovg_forest_size(const ovg_ctx *ctx);

/*
 * HIP graphs of call sequences (no counterpart in the reference).
 * ovg_graph_begin starts capturing the work that the following calls queue
 * on the context's stream; ovg_graph_end instantiates it (*n_nodes: its
 * nodes) and ovg_graph_launch replays it `repeats` times, queued without
 * waiting.  Only device work with fixed shapes can be captured: a step's
 * ovg_update_tangle + ovg_rk4_step + ovg_enforce_periodic, say.  Run the
 * sequence once before capturing it (buffers must not grow during a
 * capture: OVG_EINVAL), keep timing off, and use flows that do not depend
 * on time (the captured parameters are replayed as they are).  A call that
 * waits for the device or copies to the host ends the capture with an
 * error at ovg_graph_end.  The graph replays the same kernels on the same
 * buffers: upload nothing that changes the slot count or the topology
 * before replaying it.  ovg_graph_release frees it (ovg_destroy does too).
 */
int ovg_graph_begin(ovg_ctx *ctx);
int ovg_graph_end(ovg_ctx *ctx, int *n_nodes);
int ovg_graph_launch(ovg_ctx *ctx, int repeats);
int ovg_graph_release(ovg_ctx *ctx);

/* Upload only new positions (same n and topology). */
int ovg_upload_positions(ovg_ctx *ctx, const double *vnodes);

/* Upload only connections (same n), e.g. after a host-side topology edit. */
int ovg_upload_connections(ovg_ctx *ctx, const int *connections, const int *status);

/* Upload one per-slot double[3*n] field of the current tangle (same n):
 * the velocities used as k1 by ovg_rk4_step, or the tangents and normals
 * used by ovg_update_velocities / ovg_velocity_at. */
#define OVG_FIELD_VELS      0
#define OVG_FIELD_VS        1
#define OVG_FIELD_TANGENTS  2
#define OVG_FIELD_NORMALS   3
int ovg_upload_field(ovg_ctx *ctx, int field, const double *values);

/* Copy device state back; any pointer may be NULL.  All are double[3*n]
 * or int[2*n] in the layouts above. */
int ovg_download(ovg_ctx *ctx, double *vnodes, double *vels, double *vs,
                 double *tangents, double *normals);
int ovg_download_connections(ovg_ctx *ctx, int *connections, int *status);

int ovg_num_slots(const ovg_ctx *ctx);
/* Live (non-EMPTY) points of the current tangle. */
int ovg_num_points(const ovg_ctx *ctx);

/* ---- hot path (the reference prototypes these replace are cited) ----- */

/* update_tangents_normals, src/tangle.c:509-516 */
int ovg_update_tangents_normals(ovg_ctx *ctx);

/* update_velocities, src/tangle.c:733-749 (builds the tree when
 * method == OVG_METHOD_BH). */
int ovg_update_velocities(ovg_ctx *ctx, double t);

/* update_tangle, include/tangle.h:114-118 */
int ovg_update_tangle(ovg_ctx *ctx, double t);

/* rk4_step, src/vortex_dynamics.c:50-119.  Uses the current vels as k1,
 * exactly as the reference (the caller runs update_tangle first). */
int ovg_rk4_step(ovg_ctx *ctx, double t, double dt);

/* enforce_periodic_boundaries, src/tangle.c:793-814 */
int ovg_enforce_periodic(ovg_ctx *ctx);

/*
 * Reconnection candidates, the predicate of src/vortex_dynamics.c:197-222
 * before any reconnection of the pass: pairs (k,l), k<l, both FREE,
 * l not a neighbour of k, |seg_pwrap(x_k,x_l)| <= rec_dist.  Written as
 * k,l int pairs sorted by (k,l).  *n_pairs receives the total count even
 * when it exceeds max_pairs (then only max_pairs are written and
 * OVG_ELIMIT is returned).
 */
int ovg_reconnect_candidates(ovg_ctx *ctx, int max_pairs, int *pairs, int *n_pairs);

/*
 * reconnect, src/vortex_dynamics.c:186-288: the candidate scan on the GPU,
 * then the reference's greedy pass in (k,l) order with the angle test,
 * the approach test on direct-sum velocities and the length test.  The
 * updated connections stay on the device.  Pinned nodes are never
 * candidates (:198-212); stencils next to a pinned end step through the
 * wall as step_node does.
 */
int ovg_reconnect(ovg_ctx *ctx, double t, int *n_reconnections);

/* eliminate_small_loops, src/tangle.c:1063-1128, for closed loops: a loop
 * of m points is deleted when m - 1 < cutoff (the reference's count;
 * small_loop_cutoff = 5, src/initial_conditions.c:43).  Runs on the host
 * link mirror; after the first pass only loops through slots whose links
 * changed since the previous pass are walked, and only the deleted slots
 * are sent to the device. */
int ovg_eliminate_small_loops(ovg_ctx *ctx, int cutoff, int *n_removed);

/* remesh, src/tangle.c:1018-1061: remove a point whose forward or reverse
 * segment is shorter than min_dist when both together are shorter than
 * max_dist; insert a point on the fitted circle (add_point, :261-335) after
 * every point whose forward segment is longer than max_dist; if points were
 * added, wrap into the box and clip to the walls (ovg_enforce_walls at time
 * 0, as the reference driver calls remesh, src/main.c:131).  Lines pinned
 * to walls are taken: an open end is never removed and its stencil is
 * step_node's (src/tangle.c:155-194).  New points take the slots the reference's
 * allocator would (src/tangle.c:201-253), doubling the slot count when full
 * (expand_tangle, :68-101); ovg_num_slots reports the new count.  The
 * device flags the slots whose test (with the links before the pass) can
 * act; the host runs the reference's sequential pass over those and the
 * slots its own edits reach, and sends only the changed slots back (the
 * whole state when the slot count doubled). */
int ovg_remesh(ovg_ctx *ctx, double min_dist, double max_dist, int *n_added, int *n_removed);

/* The 22 counters of the device topology passes (OVG_TOPO_DEVICE), read
 * after the passes queued so far (waits): [0] live points, [1] error bits,
 * [2] allocator hint, [3] scratch, [4] reconnections, [5] points removed by
 * small-loop removal, [6] points added and [7] removed by remesh, [8]
 * candidate pairs taken and [9] found, [10] their endpoints, [11] rewired
 * segments (one-wavefront greedy pass), [12] endpoint target groups, [13]
 * slots the one-wavefront remesh visited, [14] endpoint velocities that
 * pass recomposed, [15] endpoints whose rewirable neighbourhood overflowed
 * its list, [16] why the last remesh left the cluster pass for the
 * one-wavefront pass (0: it did not; bits: 1 two clusters reached one
 * slot, 2 a cluster's touched slots overflowed, 4 the new-point table
 * overflowed, 8 a new point would have acted when visited, 16 more
 * candidates than the pass takes), [17] the allocator hint after the
 * cluster pass, [18] its new points, [19] remesh candidates, [20] clusters
 * of the last reconnection pass (tree-mode velocities: pairs that no other
 * cluster's rewiring reaches, one wavefront each) and [21] the candidate
 * pairs of its largest cluster.  Diagnostics, no counterpart in the
 * reference. */
int ovg_topology_counters(ovg_ctx *ctx, int *out);

/* enforce_wall_boundaries, src/tangle.c:966-1016, at time t: for every
 * solid wall (not OPEN, not PERIODIC) in face order and every FREE point in
 * slot order within rec_dist of it (check_wall, :822-840), the point's
 * tangent and direct-sum velocity (update_velocity with tree == NULL, left
 * in the context as the reference leaves them) and connect_to_wall
 * (:876-958): unless the point moves away from the wall or the wall
 * segments would be longer, it and its nearer neighbour are clipped to the
 * wall through two new nodes of status pin_mode (the reference's allocator,
 * doubling when full); then every FREE point behind a solid wall is removed.
 * The parallel test takes abs() of the cosine as an int, as the reference
 * does (:892), so only |cos| >= 1 counts.  *n_clipped / *n_removed: clips
 * and removals. */
int ovg_enforce_walls(ovg_ctx *ctx, double t, int *n_clipped, int *n_removed);

/* smooth, src/vortex_smooth.c:459-462 (low_pass_filter per loop and
 * coordinate, src/vortex_smooth.c:134-175,183-453): closed loops (those
 * spanning the box detrended and odd-extended), and open lines ending on
 * walls, which are always detrended and odd-extended and keep their ends
 * (:316-444). */
int ovg_smooth(ovg_ctx *ctx);

/*
 * Direct-sum (tree == NULL) velocity at selected slots, the single-point
 * update_velocity of src/tangle.c:600-726 as called by reconnect
 * (src/vortex_dynamics.c:250-251).  vs and vels are double[3*m] (either may
 * be NULL); tangents and normals are read from the device state.
 */
int ovg_velocity_at(ovg_ctx *ctx, int m, const int *slots, double t,
                    double *vs, double *vels);

/*
 * ovg_velocity_at without allocation (no counterpart in the reference: for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * direct sum, compose and results as ovg_velocity_at (bit for bit), but
 * d_slots (int[m]) and d_vs / d_vels (double[3*m], AoS; either may be NULL)
 * are DEVICE pointers, the work is queued on the context's stream and the
 * call returns without waiting (ovg_synchronize waits).  The chunk
 * partials and a checked copy of d_slots live in a 32 MiB device array of
 * the library, one per GPU, held until the process ends; nothing is
 * allocated or freed.  m <= OVG_NOALLOC_MAX_TARGETS, else OVG_ELIMIT.  A slot
 * out of range or EMPTY yields NaN outputs and makes the next
 * ovg_synchronize or ovg_download fail with OVG_EINVAL.  Every call on a
 * GPU shares its array: calls that overlap on one GPU (contexts driven from
 * several threads, say) race; serialise them.
 */
#define OVG_NOALLOC_MAX_TARGETS 131072
int ovg_velocity_at_noalloc(ovg_ctx *ctx, int m, const int *d_slots, double t,
                            double *d_vs, double *d_vels);

/* ---- diagnostics ----------------------------------------------------- */

/* Line length sum |seg_pwrap(x_k, x_fwd(k))| over FREE slots with a
 * forward neighbour, and the number of live points. */
int ovg_line_length(ovg_ctx *ctx, double *length, int *n_points);

/*
 * Counters of the last ovg_update_velocities: pair interactions evaluated
 * (segments and monopoles, all images), tree nodes, and the device time in
 * milliseconds of the tree build and of the Biot-Savart kernels (HIP
 * events on the context's stream).
 */
typedef struct ovg_stats {
    double   interactions;        /* pair evaluations, FP64 + FP32           */
    double   interactions_fp32;   /* the FP32 part of the above              */
    int64_t  tree_nodes;
    int32_t  tree_depth;
    int32_t  n_points;
    double   ms_tree;             /* tree build                              */
    double   ms_bs;               /* Biot-Savart kernels                     */
    double   ms_local;            /* stencil, LIA, friction, RK4 streams     */
    int32_t  bs_launches;         /* Biot-Savart kernel launches measured    */
    int32_t  pad_;
} ovg_stats;

int ovg_get_stats(ovg_ctx *ctx, ovg_stats *s);
/* Enable/disable the per-call HIP-event timing (off by default). */
int ovg_set_timing(ovg_ctx *ctx, int on);

/* Wait for all work queued on the context's stream. */
int ovg_synchronize(ovg_ctx *ctx);

/* ---- loop injection (src/vortex_injection.c) --------------------------- */
/*
 * ovg_loop_points: add_loop's geometry (src/vortex_injection.c:44-103): a
 * circle of radius r about `center` in the plane normal to `dir`, with
 * (int)(4 pi r / (dl_max + dl_min)) points, returned in *n_points; the first
 * min(cap, *n_points) go to xyz (AoS).  Host only, bit-identical to the
 * reference's points.
 * ovg_inject_loop_draw: inject_loop's schedule (:711-745) with the caller
 * holding its state: rng is the 48-bit drand48 state in erand48 form
 * (srand48(s) is {0x330E, s & 0xffff, (s >> 16) & 0xffff}) and
 * *last_injection the time of the last injection (the reference's static,
 * 0 at start).  Every call draws the direction (three draws); when
 * t - *last_injection >= 1/frequency it also draws the centre (on the top
 * face of the box) and the radius (0.05 to 0.25 box widths), sets
 * *last_injection = t and *due = 1, else *due = 0.  Host only.
 * ovg_add_loop: add_loop on a resident tangle: the points fill the next
 * empty slots in the reference's order (the slots double first while fewer
 * are empty than needed), linked in order and closed.  *n_points = points
 * added.  Loops with no point are refused (OVG_EINVAL).
 * ovg_inject_line_pairs_draw: inject_line_pairs' schedule (:760-800) with
 * the same caller-held state: when t - *last_injection >= 1/frequency, for
 * each of n_pairs pairs x1, x2 and then y1, y2 (polarized: y1 in the upper,
 * y2 in the lower half of the box's y range) into xy[4 k .. 4 k + 3] =
 * (x1, y1, x2, y2), *last_injection = t and *due = 1; else nothing is drawn
 * and *due = 0.  Host only.
 * ovg_add_line_z: add_line_in_zdir (:267-318) on a resident tangle: a
 * straight line at (x, y) from the DOWN to the UP wall (direction > 0) or
 * back, (int)(2 (zmax - zmin) / (dl_max - dl_min)) points in the next empty
 * slots (doubling first while fewer are empty), its two ends of status
 * pin_mode pinned to the walls.  A pair of the reference's injection is
 * ovg_add_line_z(x1, y1, +1) then ovg_add_line_z(x2, y2, -1).
 */
int ovg_loop_points(const double center[3], const double dir[3], double r, double dl_min,
                    double dl_max, double *xyz, int cap, int *n_points);
int ovg_inject_loop_draw(double t, double frequency, const double box_lo[3],
                         const double box_hi[3], unsigned short rng[3], double *last_injection,
                         double center[3], double dir[3], double *r, int *due);
int ovg_add_loop(ovg_ctx *ctx, const double center[3], const double dir[3], double r,
                 double dl_min, double dl_max, int *n_points);
int ovg_inject_line_pairs_draw(double t, double frequency, int n_pairs, int polarized,
                               const double box_lo[3], const double box_hi[3],
                               unsigned short rng[3], double *last_injection, double *xy,
                               int *due);
int ovg_add_line_z(ovg_ctx *ctx, double x, double y, int direction, double dl_min,
                   double dl_max, int *n_points);

/* ---- stream markers (no counterpart in the reference) ----------------- */
/*
 * Time a caller-defined sequence of calls (a time step, say) without
 * synchronising after each one: ovg_marker_record queues a marker with id
 * `id` (0 <= id < OVG_MAX_MARKERS; recording an id again moves it) on the
 * context's stream behind the work already queued; ovg_marker_elapsed waits
 * for marker id1 and returns the device time in milliseconds between the
 * two markers (both recorded, id0 before id1).
 */
#define OVG_MAX_MARKERS 65536
int ovg_marker_record(ovg_ctx *ctx, int id);
int ovg_marker_elapsed(ovg_ctx *ctx, int id0, int id1, double *ms);

/* ---- autotuning (no counterpart in the reference) --------------------- */
/*
 * Time the list traversal (OVG_TRAVERSE_LIST) on the context's current
 * tangle (or forest), one warm-up and `repeats` timed launches each on a
 * tree built once, and keep the fastest in the context's parameters.  Stage
 * one: every combination of points_per_wave {8, 16, 21, 32, 64}, leaf_size
 * {1, 2, 4, 8} and group_level {L-1, L, L+1} (L: the automatic level) at
 * the context's launch shape.  Stage two, at the best of stage one: the
 * launch shape, list_stack {128, 256, 384} x list_tile {96, 128} x
 * list_blocks {0, 8, 16} (traversal stack depth, far-field flush tile,
 * persistent grid).  Results differ between configurations only within the
 * traversal's truncation (each is never coarser than the reference's
 * per-point acceptance) and the FP32 summation order.
 */
typedef struct ovg_tune_row {
    int32_t points_per_wave, leaf_size, group_level;
    int32_t list_stack, list_tile, list_blocks;  /* the launch shape (ovg_params) */
    double  ms;             /* average launch, HIP events               */
    double  pairs;          /* pair interactions per launch            */
} ovg_tune_row;
#define OVG_TUNE_MAX_ROWS 96
typedef struct ovg_tune {
    int32_t n_rows, best;
    ovg_tune_row rows[OVG_TUNE_MAX_ROWS];
} ovg_tune;
int ovg_autotune(ovg_ctx *ctx, int repeats, ovg_tune *out);
size_t ovg_tune_sizeof(void);

/* ---- tangle files (src/utils.c:68-199) ------------------------------- */
/*
 * One text line per live point, 18 tab-separated columns: loop index,
 * position, vels, tangents, normals (3 each), slot, reverse, forward,
 * status, pin_wall; loops in the slot order of their first point
 * (save_tangle, src/utils.c:132-184).  digits = 15 writes the reference's
 * %.15g byte for byte; 17 round-trips every double exactly.  vels,
 * tangents, normals may be NULL (written as zeros).
 */
int ovg_write_tangle_file(const char *path, int n, const double *vnodes, const double *vels,
                          const double *tangents, const double *normals,
                          const int *connections, const int *status, int digits);
/*
 * Read a tangle file (load_tangle / load_point, src/utils.c:93-130,186-199)
 * into slot arrays: each line's point goes to its slot; the slot count is
 * 512 doubled until every slot id fits (create_tangle(512) and
 * expand_tangle, src/main.c:85, src/utils.c:117-119); slots not in the file
 * are EMPTY with zero vectors.  Call with all arrays NULL to get the slot
 * count in *n_slots, then with arrays of that many slots.
 */
int ovg_read_tangle_file(const char *path, int n_cap, int *n_slots, double *vnodes,
                         double *vels, double *tangents, double *normals, int *connections,
                         int *status);
/* The context's tangle to a file / a file into the context (replacing its
 * tangle; vels, tangents and normals as in the file). */
int ovg_save_tangle(ovg_ctx *ctx, const char *path, int digits);
int ovg_load_tangle(ovg_ctx *ctx, const char *path);

/* ---- seeded tangle generator (host code, no GPU) --------------------- */
/*
 * Rings with radius ~ U[r_min, r_max], centre ~ U[0,D)^3 and normal
 * uniform on the sphere, n = floor(2 pi R / dxi) points each (the point
 * count of add_loop, src/vortex_injection.c:52, with
 * dl_min + dl_max = 2 dxi), positions wrapped into [0,D)^3.  Ring k uses
 * the SplitMix64 stream of (seed, k).  Slots are filled contiguously from
 * 0; the remaining capacity (n_cap) is EMPTY.  Returns the number of live
 * points, or -1 when it would exceed n_cap.  Any output may be NULL to
 * only count.
 */
int ovg_generate_rings(uint64_t seed, int n_rings, double D, double r_min,
                       double r_max, double dxi, int n_cap, double *vnodes,
                       int *connections, int *status);

/* Number of rings needed to reach at least `target_points` with the same
 * stream as ovg_generate_rings. */
int ovg_rings_for_points(uint64_t seed, int target_points, double D,
                         double r_min, double r_max, double dxi);

#ifdef __cplusplus
}
#endif

#endif /* OPENVORT_GPU_H */
