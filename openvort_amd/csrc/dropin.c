/*
 * dropin.c -- the reference's hot-path prototypes (include/openvort_dropin.h)
 * on top of the GPU C ABI (include/openvort_gpu.h).
 *
 * The reference globals below are defined by the reference's
 * src/initial_conditions.c in the program that links this library.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "openvort_dropin.h"
#include "openvort_gpu.h"

/* include/vortex_constants.h:11-60 */
extern double VORTEX_WIDTH, KAPPA;
extern int global_num_threads;
extern int use_mutual_friction;
extern double alpha, alpha_p;
extern double reconnection_angle_cutoff, rec_dist;
extern double freq_to_cutoff;
extern double global_dl_min, global_dl_max; /* include/vortex_constants.h:33-34 */
extern int use_BH;
extern double BH_resolution;

/* include/external_velocity.h:11-19,29-33 */
struct ext_vel_param {
    int type;
    double strength;
    double time_of_linear_decrease;
    double starting_strength;
    struct vec3 direction;
};
extern struct ext_vel_param vn_conf, vs_conf, vb_conf;

static ovg_ctx *g_ctx;
static int g_precision = OVG_PREC_FP64;
static int g_traverse = OVG_TRAVERSE_POINT;
static int g_leaf = 1;
static int g_configured;

static void die(const char *where, int code)
{
    fprintf(stderr, "openvort_dropin: %s failed (code %d): %s\n", where, code, ovg_last_error());
    abort();
}

#define CK(call)                        \
    do {                                \
        int rc_ = (call);               \
        if (rc_ != OVG_OK) die(#call, rc_); \
    } while (0)

void openvort_dropin_configure(int precision, int traverse, int leaf_size)
{
    g_precision = precision;
    g_traverse = traverse;
    g_leaf = leaf_size;
    g_configured = 1;
}

static void read_env(void)
{
    if (g_configured) return;
    const char *s = getenv("OPENVORT_PRECISION");
    if (s && strcmp(s, "mixed") == 0) g_precision = OVG_PREC_MIXED;
    s = getenv("OPENVORT_LEAF");
    if (s) g_leaf = atoi(s);
    /* the point traversal is the reference's FP64 split-until-one walk */
    if (g_precision != OVG_PREC_FP64 || g_leaf != 1) g_traverse = OVG_TRAVERSE_LIST;
    s = getenv("OPENVORT_TRAVERSE");
    if (s && strcmp(s, "group") == 0) g_traverse = OVG_TRAVERSE_GROUP;
    if (s && strcmp(s, "list") == 0) g_traverse = OVG_TRAVERSE_LIST;
    g_configured = 1;
}

static void set_flow(ovg_flow *d, const struct ext_vel_param *s)
{
    d->type = s->type;
    d->strength = s->strength;
    d->time_of_linear_decrease = s->time_of_linear_decrease;
    d->starting_strength = s->starting_strength;
    for (int k = 0; k < 3; ++k) d->direction[k] = s->direction.p[k];
}

/* Parameters from the reference globals and the tangle's box and image
 * table; then upload the tangle. */
static ovg_ctx *prepare(const struct tangle_state *t)
{
    read_env();
    if (!g_ctx) CK(ovg_create(&g_ctx, 0));
    ovg_params p;
    ovg_params_default(&p, 1.0);
    p.kappa = KAPPA;
    p.vortex_width = VORTEX_WIDTH;
    p.use_mutual_friction = use_mutual_friction;
    p.alpha = alpha;
    p.alpha_p = alpha_p;
    set_flow(&p.vn, &vn_conf);
    set_flow(&p.vs, &vs_conf);
    set_flow(&p.vb, &vb_conf);
    for (int k = 0; k < 3; ++k) {
        p.box_lo[k] = t->box.bottom_left_back.p[k];
        p.box_hi[k] = t->box.top_right_front.p[k];
    }
    for (int f = 0; f < 6; ++f) p.wall[f] = (int)t->box.wall[f];
    if (t->bimg.n > 26) die("image table larger than 26", OVG_EINVAL);
    p.n_images = t->bimg.n;
    for (int j = 0; j < t->bimg.n; ++j) {
        const struct image_tangle *im = &t->bimg.images[j];
        for (int k = 0; k < 3; ++k) p.images[j][k] = im->shift[k];
        /* the faces of reflect[0 .. number_of_reflects) (src/tangle.c:381-385) */
        p.image_reflect[j] = 0;
        for (int r = 0; r < im->number_of_reflects && r < 3; ++r)
            if (im->reflect[r] >= 0 && im->reflect[r] < 6) p.image_reflect[j] |= 1 << im->reflect[r];
    }
    p.method = use_BH ? OVG_METHOD_BH : OVG_METHOD_DIRECT;
    p.theta = BH_resolution;
    p.precision = g_precision;
    /* mixed precision: pairs within two of the largest point spacings stay FP64 */
    p.mixed_near = 2 * global_dl_max;
    p.traverse = g_traverse;
    p.leaf_size = g_leaf;
    p.rec_dist = rec_dist;
    p.rec_angle = reconnection_angle_cutoff;
    p.freq_to_cutoff = freq_to_cutoff;
    CK(ovg_set_params(g_ctx, &p));
    CK(ovg_upload(g_ctx, t->N, (const double *)t->vnodes, (const int *)t->connections,
                  (const int *)t->status));
    return g_ctx;
}

void update_tangents_normals(struct tangle_state *tangle)
{
    ovg_ctx *c = prepare(tangle);
    CK(ovg_update_tangents_normals(c));
    CK(ovg_download(c, NULL, NULL, NULL, (double *)tangle->tangents, (double *)tangle->normals));
}

/* update_velocities reads the caller's tangents and normals as they are
 * (src/tangle.c:588,707-713), so they are uploaded with the positions. */
void update_velocities(struct tangle_state *tangle, double t)
{
    ovg_ctx *c = prepare(tangle);
    CK(ovg_upload_field(c, OVG_FIELD_TANGENTS, (const double *)tangle->tangents));
    CK(ovg_upload_field(c, OVG_FIELD_NORMALS, (const double *)tangle->normals));
    CK(ovg_update_velocities(c, t));
    CK(ovg_download(c, NULL, (double *)tangle->vels, (double *)tangle->vs, NULL, NULL));
}

void update_velocity(struct tangle_state *tangle, int k, double t, struct octree *tree)
{
    (void)tree;
    if (tangle->status[k].status == EMPTY) return;
    ovg_ctx *c = prepare(tangle);
    CK(ovg_upload_field(c, OVG_FIELD_TANGENTS, (const double *)tangle->tangents));
    CK(ovg_upload_field(c, OVG_FIELD_NORMALS, (const double *)tangle->normals));
    CK(ovg_velocity_at(c, 1, &k, t, tangle->vs[k].p, tangle->vels[k].p));
}

/* k1 is the caller's vels, from its preceding update_tangle
 * (src/main.c:135-136, src/vortex_dynamics.c:102). */
void rk4_step(struct tangle_state *tangle, double t, double dt)
{
    ovg_ctx *c = prepare(tangle);
    CK(ovg_upload_field(c, OVG_FIELD_VELS, (const double *)tangle->vels));
    CK(ovg_rk4_step(c, t, dt));
    CK(ovg_download(c, (double *)tangle->vnodes, NULL, NULL, NULL, NULL));
}

int reconnect(struct tangle_state *tangle, double t)
{
    ovg_ctx *c = prepare(tangle);
    int n = 0;
    CK(ovg_reconnect(c, t, &n));
    CK(ovg_download_connections(c, (int *)tangle->connections, NULL));
    for (int k = 0; k < tangle->N; ++k) tangle->recalculate[k] = 0;
    return n;
}

void smooth(struct tangle_state *tangle)
{
    ovg_ctx *c = prepare(tangle);
    CK(ovg_smooth(c));
    CK(ovg_download(c, (double *)tangle->vnodes, NULL, NULL, NULL, NULL));
}

void enforce_periodic_boundaries(struct tangle_state *tangle)
{
    ovg_ctx *c = prepare(tangle);
    CK(ovg_enforce_periodic(c));
    CK(ovg_download(c, (double *)tangle->vnodes, NULL, NULL, NULL, NULL));
}
