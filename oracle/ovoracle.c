/*
 * ovoracle.c -- CPU oracle (test infrastructure only; see ovoracle.h).
 *
 * Restates the reference FP64 algorithm in the reference's operation
 * order.  Build with -O2 -fno-fast-math -ffp-contract=off (oracle/Makefile)
 * so that every expression rounds as written.
 */
#include "ovoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_E
#define M_E 2.7182818284590452354
#endif

enum { LEFT = 0, RIGHT, BACK, FRONT, DOWN, UP };

/* ---- vector helpers with the reference's operation order ------------- */
/* src/vec3_math.c:91-96: ((0 + u0 v0) + u1 v1) + u2 v2 */
static double dot3(const double *u, const double *v)
{
    double out = 0;
    for (int k = 0; k < 3; ++k) out += u[k] * v[k];
    return out;
}
/* src/vec3_math.c:167-174 */
static double len3(const double *u)
{
    double out = 0;
    for (int k = 0; k < 3; ++k) out += u[k] * u[k];
    return sqrt(out);
}
/* src/vec3_math.c:182-187: |u - v| */
static double dist3(const double *u, const double *v)
{
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = u[k] - v[k];
    return len3(d);
}
/* src/vec3_math.c:104-111 */
static double ndot3(const double *u, const double *v)
{
    double dot = dot3(u, v);
    double d1 = len3(u);
    double d2 = len3(v);
    return dot / d1 / d2;
}
/* src/vec3_math.c:119-127 */
static void cross3(double *res, const double *u_, const double *v_)
{
    double u[3] = {u_[0], u_[1], u_[2]}, v[3] = {v_[0], v_[1], v_[2]};
    res[0] = u[1] * v[2] - u[2] * v[1];
    res[1] = u[2] * v[0] - u[0] * v[2];
    res[2] = u[0] * v[1] - u[1] * v[0];
}

/* seg_pwrap, src/boundaries.c:54-83: r2 moved by whole periods until it
 * is within half a box of r1 on each periodic axis. */
static void pwrap(const double *r1, const double *r2, const ovg_params *p,
                  double *o)
{
    o[0] = r2[0]; o[1] = r2[1]; o[2] = r2[2];
    double Lx = p->box_hi[0] - p->box_lo[0];
    double Ly = p->box_hi[1] - p->box_lo[1];
    double Lz = p->box_hi[2] - p->box_lo[2];
    if (p->wall[LEFT] == OVG_WALL_PERIODIC) {
        while ((o[0] - r1[0]) > Lx / 2) o[0] -= Lx;
        while ((o[0] - r1[0]) < -Lx / 2) o[0] += Lx;
    }
    if (p->wall[BACK] == OVG_WALL_PERIODIC) {
        while ((o[1] - r1[1]) > Ly / 2) o[1] -= Ly;
        while ((o[1] - r1[1]) < -Ly / 2) o[1] += Ly;
    }
    if (p->wall[DOWN] == OVG_WALL_PERIODIC) {
        while ((o[2] - r1[2]) > Lz / 2) o[2] -= Lz;
        while ((o[2] - r1[2]) < -Lz / 2) o[2] += Lz;
    }
}

/* segment_field body, src/tangle.c:536-559: seg.r1 = r1, seg.r2 = r2. */
void ovo_segment_field(const double r1[3], const double r2[3],
                       const double r[3], double kappa, double out[3])
{
    double R[3], Rp1[3];
    for (int k = 0; k < 3; ++k) R[k] = r1[k] - r[k];
    for (int k = 0; k < 3; ++k) Rp1[k] = r2[k] - r[k];
    double lR = len3(R);
    double lRp1 = len3(Rp1);
    double denom = lR * lRp1 * (lR * lRp1 + dot3(R, Rp1));
    double f = kappa / 4 / M_PI;
    out[0] = out[1] = out[2] = 0;
    if (lR < 1e-8 || lRp1 < 1e-8) return;
    if (ndot3(R, Rp1) < 1e-8) return;
    double vv[3];
    cross3(vv, R, Rp1);
    double m = f * (lR + lRp1) / denom;
    for (int k = 0; k < 3; ++k) out[k] = vv[k] * m;
}

/* segment_field, src/tangle.c:525-560 */
static void segment_field(const double *vnodes, const int *conn,
                          const ovg_params *p, int i, const double *r,
                          double *out)
{
    int next = conn[2 * i];
    out[0] = out[1] = out[2] = 0;
    if (next == -1) return;
    double r2[3];
    pwrap(vnodes + 3 * i, vnodes + 3 * next, p, r2);
    ovo_segment_field(vnodes + 3 * i, r2, r, p->kappa, out);
}

/* mirror_shift, src/boundaries.c:159-193: the coordinate of the wall's
 * axis becomes 2 * wall - v */
static void mirror_shift(const double *v, int wall, const ovg_params *p, double *out)
{
    for (int c = 0; c < 3; ++c) out[c] = v[c];
    int a = wall / 2;
    double wall_pos = wall % 2 ? p->box_hi[a] : p->box_lo[a];
    out[a] = 2 * wall_pos - out[a];
}

/* step_node, src/tangle.c:155-194, into out: FREE nodes walk the links; a
 * pinned node (status NULL = all FREE) at its open end walks the other way
 * and mirrors the result through its wall. */
static void step_node(const double *vnodes, const int *conn, const int *status,
                      const ovg_params *p, int i, int where, double *out)
{
    if (where == 0) {
        for (int c = 0; c < 3; ++c) out[c] = vnodes[3 * i + c];
        return;
    }
    int st = status ? status[2 * i] : OVG_FREE;
    if (st == OVG_PINNED || st == OVG_PINNED_SLIP) {
        int next = where > 0 ? conn[2 * i] : conn[2 * i + 1];
        if (next >= 0) {
            step_node(vnodes, conn, status, p, next, where > 0 ? where - 1 : where + 1, out);
            return;
        }
        double o[3];
        step_node(vnodes, conn, status, p, i, -where, o);
        mirror_shift(o, status[2 * i + 1], p, out);
        return;
    }
    step_node(vnodes, conn, status, p, where > 0 ? conn[2 * i] : conn[2 * i + 1],
              where > 0 ? where - 1 : where + 1, out);
}

/* update_tangent_normal, src/tangle.c:423-502 */
static void tangent_normal(const double *vnodes, const int *conn, const int *status,
                           const ovg_params *p, int k, double *tan,
                           double *nor)
{
    const double *s0 = vnodes + 3 * k;
    double s1[3], s2[3], sm1[3], sm2[3];
    step_node(vnodes, conn, status, p, k, 1, s1);
    step_node(vnodes, conn, status, p, k, 2, s2);
    step_node(vnodes, conn, status, p, k, -1, sm1);
    step_node(vnodes, conn, status, p, k, -2, sm2);
    double w[4][3], w12[3], wm12[3];
    pwrap(s0, s2, p, w[0]);
    pwrap(s0, s1, p, w[1]);
    pwrap(s0, sm1, p, w[2]);
    pwrap(s0, sm2, p, w[3]);
    pwrap(s1, s2, p, w12);
    pwrap(sm1, sm2, p, wm12);
    double ds[4][3];
    for (int j = 0; j < 4; ++j)
        for (int c = 0; c < 3; ++c) ds[j][c] = w[j][c] - s0[c];
    double d1 = dist3(s0, w[1]);
    double d2 = d1 + dist3(s1, w12);
    double dm1 = dist3(s0, w[2]);
    double dm2 = dm1 + dist3(sm1, wm12);
    double dsd[4] = {
        d2 * (d2 - d1) * (dm1 + d2) * (dm2 + d2),
        d1 * (d2 - d1) * (dm1 + d1) * (dm2 + d1),
        dm1 * (dm1 + d1) * (dm1 + d2) * (dm2 - dm1),
        dm2 * (dm2 + d1) * (dm2 + d2) * (dm2 - dm1)};
    double c1[4] = {-d1 * dm1 * dm2, d2 * dm1 * dm2, -d1 * d2 * dm2,
                    d1 * d2 * dm1};
    double c2[4] = {2 * ((dm1 - d1) * dm2 - d1 * dm1),
                    -2 * ((dm1 - d2) * dm2 - d2 * dm1),
                    2 * ((d2 + d1) * dm2 - d1 * d2),
                    -2 * ((d2 + d1) * dm1 - d1 * d2)};
    for (int i = 0; i < 3; ++i) {
        tan[i] = 0;
        nor[i] = 0;
        for (int z = 0; z < 4; ++z) {
            tan[i] += c1[z] / dsd[z] * ds[z][i];
            nor[i] += c2[z] / dsd[z] * ds[z][i];
        }
    }
}

void ovo_tangents_normals(int n, const double *vnodes, const int *conn,
                          const int *status, const ovg_params *p,
                          double *tangents, double *normals)
{
    for (int k = 0; k < n; ++k) {
        if (status[2 * k] == OVG_EMPTY) continue;
        tangent_normal(vnodes, conn, status, p, k, tangents + 3 * k, normals + 3 * k);
    }
}

/* get_v, src/external_velocity.c:13-81 (DECREASING included). */
static void ext_flow(const ovg_flow *f, double t, double *res)
{
    res[0] = res[1] = res[2] = 0;
    if (f->type == OVG_FLOW_NONE) return;
    double str = f->strength;
    if (f->type == OVG_FLOW_DECREASING) {
        if (t < f->time_of_linear_decrease)
            str = f->starting_strength -
                  t * (f->starting_strength - f->strength) /
                      f->time_of_linear_decrease;
        else
            str = f->strength;
    }
    double d[3] = {f->direction[0], f->direction[1], f->direction[2]};
    if (!(d[0] == 0 && d[1] == 0 && d[2] == 0)) {
        double l = len3(d);
        for (int k = 0; k < 3; ++k) d[k] = d[k] * (1 / l);
    }
    for (int k = 0; k < 3; ++k) res[k] = d[k] * str;
}

/* lia_velocity, src/tangle.c:568-592 */
static void lia(const double *vnodes, const int *conn, const int *status,
                const ovg_params *p, int i, const double *tan, const double *nor,
                double *out)
{
    const double *x = vnodes + 3 * i;
    double next[3], prev[3];
    step_node(vnodes, conn, status, p, i, 1, next);
    step_node(vnodes, conn, status, p, i, -1, prev);
    double wf[3], wr[3];
    pwrap(x, next, p, wf);
    pwrap(prev, x, p, wr);
    double l_next = dist3(x, wf);
    double l_prev = dist3(prev, wr);
    double f = p->kappa *
               log(2 * sqrt(l_next * l_prev) / sqrt(M_E) / p->vortex_width) /
               4 / M_PI;
    double vv[3];
    cross3(vv, tan, nor);
    for (int k = 0; k < 3; ++k) out[k] = vv[k] * f;
}

/* shifted, src/tangle.c:370-414: the periodic shift, then each mirror
 * face of image j (bit f of image_reflect[j], faces LEFT RIGHT BACK FRONT
 * DOWN UP) sets its coordinate to 2 * wall - r (:385-400) */
static void shifted(const double *r, int j, const ovg_params *p, double *rs)
{
    for (int k = 0; k < 3; ++k) {
        double L = p->box_hi[k] - p->box_lo[k];
        rs[k] = r[k];
        rs[k] -= L * p->images[j][k];
    }
    for (int f = 0; f < 6; ++f) {
        if (!(p->image_reflect[j] >> f & 1)) continue;
        int a = f / 2;
        double wall = f % 2 ? p->box_hi[a] : p->box_lo[a];
        rs[a] = 2 * wall - r[a];
    }
}

/* mirror_vector_reflect, src/boundaries.c:201-211, for every mirror face of
 * image j, with boundary_normals of src/boundaries.c:11-18 */
static void mirror_reflect(double *v, int j, const ovg_params *p)
{
    for (int f = 0; f < 6; ++f) {
        if (!(p->image_reflect[j] >> f & 1)) continue;
        double n[3] = {0, 0, 0};
        n[f / 2] = f % 2 ? -1 : 1;
        double nc = 0;
        for (int c = 0; c < 3; ++c) nc += v[c] * n[c];
        double tmp[3];
        for (int c = 0; c < 3; ++c) tmp[c] = n[c] * (-2 * nc);
        for (int c = 0; c < 3; ++c) v[c] = v[c] + tmp[c];
    }
}

/* ---------------- reference octree (src/octree.c) ---------------------- */
typedef struct rnode {
    int N;            /* node count in this cell                          */
    int s;            /* start in the id pool                              */
    int has_children; /* internal cell: 8 children exist                   */
    int child[8];
    double com[3], circ[3];
    double lo[3], hi[3];
} rnode;

typedef struct rtree {
    rnode *nodes;
    long n_nodes, cap;
    int *pool;       /* node ids, partitioned per node                      */
    int *tmp;
    int depth;
} rtree;

static int rtree_new(rtree *t)
{
    if (t->n_nodes == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->nodes = (rnode *)realloc(t->nodes, sizeof(rnode) * t->cap);
    }
    rnode *nd = &t->nodes[t->n_nodes];
    memset(nd, 0, sizeof(*nd));
    for (int c = 0; c < 8; ++c) nd->child[c] = -1;
    return (int)t->n_nodes++;
}

/* in_octree_box, src/octree.c:88-103 (closed bounds) */
static int in_cell(const double *lo, const double *hi, const double *v)
{
    return v[0] >= lo[0] && v[0] <= hi[0] && v[1] >= lo[1] && v[1] <= hi[1] &&
           v[2] >= lo[2] && v[2] <= hi[2];
}

/* make_child_box, src/octree.c:165-232: the exact bound arithmetic of each
 * of the eight children (BLF BLB BRF BRB TLF TLB TRF TRB). */
static void child_box(const double *plo, const double *phi, int idx,
                      double *lo, double *hi)
{
    double c[3], L[3];
    for (int k = 0; k < 3; ++k) c[k] = plo[k] + phi[k];
    for (int k = 0; k < 3; ++k) c[k] = c[k] * 0.5;
    for (int k = 0; k < 3; ++k) L[k] = phi[k] - plo[k];
    switch (idx) {
    case 0: /* BLF */
        memcpy(lo, plo, 24); lo[0] += L[0] / 2;
        memcpy(hi, c, 24); hi[0] += L[0] / 2;
        break;
    case 1: /* BLB */
        memcpy(lo, plo, 24); memcpy(hi, c, 24);
        break;
    case 2: /* BRF */
        memcpy(lo, c, 24); lo[2] -= L[2] / 2;
        memcpy(hi, phi, 24); hi[2] -= L[2] / 2;
        break;
    case 3: /* BRB */
        memcpy(lo, plo, 24); lo[1] += L[1] / 2;
        memcpy(hi, c, 24); hi[1] += L[1] / 2;
        break;
    case 4: /* TLF */
        memcpy(lo, c, 24); lo[1] -= L[1] / 2;
        memcpy(hi, phi, 24); hi[1] -= L[1] / 2;
        break;
    case 5: /* TLB */
        memcpy(lo, plo, 24); lo[2] += L[2] / 2;
        memcpy(hi, c, 24); hi[2] += L[2] / 2;
        break;
    case 6: /* TRF */
        memcpy(lo, c, 24); memcpy(hi, phi, 24);
        break;
    default: /* TRB */
        memcpy(lo, c, 24); lo[0] -= L[0] / 2;
        memcpy(hi, phi, 24); hi[0] -= L[0] / 2;
        break;
    }
}

/* octree_inner_sort, src/octree.c:110-132: ids move, in order, to the
 * first child whose closed box holds them; ids in no child are dropped
 * from the subtree (they stay counted in this cell). */
static void node_means(rtree *t, int ni, const double *vnodes,
                       const int *conn, const ovg_params *p);

static void rtree_sort(rtree *t, int ni, const double *vnodes,
                       const int *conn, const ovg_params *p, int level)
{
    if (level > t->depth) t->depth = level;
    /* the cell's ids are still in its node_ids order here */
    node_means(t, ni, vnodes, conn, p);
    rnode nd = t->nodes[ni];
    if (nd.N <= 1) return;
    int kids[8];
    for (int c = 0; c < 8; ++c) {
        kids[c] = rtree_new(t);
        rnode *ch = &t->nodes[kids[c]];
        rnode *par = &t->nodes[ni];
        child_box(par->lo, par->hi, c, ch->lo, ch->hi);
    }
    rnode *par = &t->nodes[ni];
    par->has_children = 1;
    for (int c = 0; c < 8; ++c) par->child[c] = kids[c];
    /* stable partition of pool[s, s+N) into the children, in child order */
    int cnt[9] = {0};
    int *which = (int *)malloc(sizeof(int) * (size_t)nd.N);
    for (int k = 0; k < nd.N; ++k) {
        const double *v = vnodes + 3 * t->pool[nd.s + k];
        int w = 8;
        for (int c = 0; c < 8; ++c) {
            rnode *ch = &t->nodes[kids[c]];
            if (in_cell(ch->lo, ch->hi, v)) { w = c; break; }
        }
        which[k] = w;
        cnt[w]++;
    }
    int off[9], acc = nd.s;
    for (int c = 0; c < 9; ++c) { off[c] = acc; acc += cnt[c]; }
    for (int c = 0; c < 8; ++c) {
        t->nodes[kids[c]].s = off[c];
        t->nodes[kids[c]].N = cnt[c];
    }
    for (int k = 0; k < nd.N; ++k) t->tmp[off[which[k]]++] = t->pool[nd.s + k];
    memcpy(t->pool + nd.s, t->tmp + nd.s, sizeof(int) * (size_t)nd.N);
    free(which);
    for (int c = 0; c < 8; ++c) rtree_sort(t, kids[c], vnodes, conn, p, level + 1);
}

/* octree_update_means, src/octree.c:258-286, for one cell: its ids are
 * summed in node_ids order (slot order filtered down the tree). */
static void node_means(rtree *t, int ni, const double *vnodes,
                       const int *conn, const ovg_params *p)
{
    rnode *nd = &t->nodes[ni];
    if (nd->N == 0) return;
    double com[3] = {0, 0, 0};
    for (int k = 0; k < nd->N; ++k) {
        const double *v = vnodes + 3 * t->pool[nd->s + k];
        for (int c = 0; c < 3; ++c) com[c] = com[c] + v[c];
    }
    for (int c = 0; c < 3; ++c) com[c] = com[c] * (1.0 / nd->N);
    double circ[3] = {0, 0, 0};
    for (int k = 0; k < nd->N; ++k) {
        int idx = t->pool[nd->s + k];
        int fwd = conn[2 * idx];
        if (fwd > -1) {
            double w[3];
            pwrap(vnodes + 3 * idx, vnodes + 3 * fwd, p, w);
            for (int c = 0; c < 3; ++c)
                circ[c] = circ[c] + (w[c] - vnodes[3 * idx + c]);
        }
    }
    memcpy(nd->com, com, 24);
    memcpy(nd->circ, circ, 24);
}

/* octree_build, src/octree.c:139-156 */
static void rtree_build(rtree *t, int n, const double *vnodes, const int *conn,
                        const int *status, const ovg_params *p)
{
    memset(t, 0, sizeof(*t));
    t->pool = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    t->tmp = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int root = rtree_new(t);
    int cnt = 0;
    for (int k = 0; k < n; ++k)
        if (status[2 * k] != OVG_EMPTY) t->pool[cnt++] = k;
    rnode *r = &t->nodes[root];
    r->N = cnt;
    r->s = 0;
    memcpy(r->lo, p->box_lo, 24);
    memcpy(r->hi, p->box_hi, 24);
    rtree_sort(t, root, vnodes, conn, p, 0);
}

static void rtree_free(rtree *t)
{
    free(t->nodes);
    free(t->pool);
    free(t->tmp);
}

/* octree_get_vs, src/octree.c:296-356 */
static void rtree_vs(const rtree *t, int ni, const double *r, double theta,
                     double kappa, double *res, double *count)
{
    const rnode *nd = &t->nodes[ni];
    res[0] = res[1] = res[2] = 0;
    if (nd->N == 0) return;
    double Lx = nd->hi[0] - nd->lo[0];
    double Ly = nd->hi[1] - nd->lo[1];
    double Lz = nd->hi[2] - nd->lo[2];
    double Lm = Lx > Ly ? (Lx > Lz ? Lx : Lz) : (Ly > Lz ? Ly : Lz);
    double d = dist3(r, nd->com);
    if (d < 1e-8) return;
    if (Lm / d < theta || nd->N == 1) {
        double R[3], Rp1[3];
        for (int k = 0; k < 3; ++k) R[k] = nd->com[k] - r[k];
        for (int k = 0; k < 3; ++k) Rp1[k] = R[k] + nd->circ[k];
        double lR = len3(R);
        double lRp1 = len3(Rp1);
        double denom = lR * lRp1 * (lR * lRp1 + dot3(R, Rp1));
        double f = kappa / 4 / M_PI;
        *count += 1;
        if (lR < 1e-8 || lRp1 < 1e-8) return;
        if (ndot3(R, Rp1) < 1e-8) return;
        double vv[3];
        cross3(vv, R, Rp1);
        double m = f * (lR + lRp1) / denom;
        for (int k = 0; k < 3; ++k) res[k] = vv[k] * m;
        return;
    }
    double total[3] = {0, 0, 0}, part[3];
    for (int kk = 0; kk < 8; ++kk) {
        rtree_vs(t, nd->child[kk], r, theta, kappa, part, count);
        for (int c = 0; c < 3; ++c) total[c] = total[c] + part[c];
    }
    memcpy(res, total, 24);
}

void ovo_bh_tree_stats(int n, const double *vnodes, const int *conn,
                       const int *status, const ovg_params *p, long *nodes,
                       int *depth, long *nonempty)
{
    rtree t;
    rtree_build(&t, n, vnodes, conn, status, p);
    long ne = 0;
    for (long i = 0; i < t.n_nodes; ++i) ne += t.nodes[i].N > 0;
    if (nodes) *nodes = t.n_nodes;
    if (depth) *depth = t.depth;
    if (nonempty) *nonempty = ne;
    rtree_free(&t);
}

/* update_velocity, src/tangle.c:600-726, FREE nodes (PINNED gets
 * vs = 0, vels = v_b as the reference).  tree == NULL selects the direct
 * sum. */
static void velocity_one(int n, const double *vnodes, const int *conn,
                         const int *status, const ovg_params *p, double t,
                         const double *tangents, const double *normals,
                         const rtree *tree, int k, double *vs, double *vels,
                         double *count)
{
    if (status[2 * k] == OVG_EMPTY) return;
    if (status[2 * k] == OVG_PINNED) {
        vs[0] = vs[1] = vs[2] = 0;
        ext_flow(&p->vb, t, vels);
        return;
    }
    const double *x = vnodes + 3 * k;
    lia(vnodes, conn, status, p, k, tangents + 3 * k, normals + 3 * k, vs);
    double evs[3];
    ext_flow(&p->vs, t, evs);
    for (int c = 0; c < 3; ++c) vs[c] = vs[c] + evs[c];
    double tot[3] = {0, 0, 0};
    if (tree) {
        double vt[3];
        rtree_vs(tree, 0, x, p->theta, p->kappa, vt, count);
        for (int c = 0; c < 3; ++c) vs[c] = vs[c] + vt[c];
        for (int j = 0; j < p->n_images; ++j) {
            double rs[3], vsh[3];
            shifted(x, j, p, rs);
            rtree_vs(tree, 0, rs, p->theta, p->kappa, vsh, count);
            mirror_reflect(vsh, j, p);  /* src/tangle.c:644-646 */
            for (int c = 0; c < 3; ++c) tot[c] = tot[c] + vsh[c];
        }
    } else {
        for (int m = 0; m < n; ++m) {
            if (conn[2 * m] == -1 || k == m || k == conn[2 * m]) continue;
            double sv[3];
            segment_field(vnodes, conn, p, m, x, sv);
            *count += 1;
            for (int c = 0; c < 3; ++c) vs[c] = vs[c] + sv[c];
        }
        for (int j = 0; j < p->n_images; ++j) {
            double rs[3];
            shifted(x, j, p, rs);
            for (int m = 0; m < n; ++m) {
                if (conn[2 * m] == -1) continue;
                double sv[3];
                segment_field(vnodes, conn, p, m, rs, sv);
                *count += 1;
                for (int c = 0; c < 3; ++c) vs[c] = vs[c] + sv[c];
            }
        }
    }
    for (int c = 0; c < 3; ++c) vs[c] = vs[c] + tot[c];
    for (int c = 0; c < 3; ++c) vels[c] = vs[c];
    if (p->use_mutual_friction) {
        double dv[3], tmp[3];
        const double *tn = tangents + 3 * k;
        ext_flow(&p->vn, t, dv);
        for (int c = 0; c < 3; ++c) dv[c] = dv[c] - vs[c];
        cross3(tmp, tn, dv);
        for (int c = 0; c < 3; ++c) tmp[c] = tmp[c] * p->alpha;
        for (int c = 0; c < 3; ++c) vels[c] = vels[c] + tmp[c];
        cross3(tmp, tn, dv);
        cross3(tmp, tn, tmp);
        for (int c = 0; c < 3; ++c) tmp[c] = tmp[c] * -p->alpha_p;
        for (int c = 0; c < 3; ++c) vels[c] = vels[c] + tmp[c];
    }
    if (status[2 * k] == OVG_PINNED_SLIP) {  /* src/tangle.c:718-725 */
        int w = status[2 * k + 1];
        double nrm[3] = {0, 0, 0};
        nrm[w / 2] = w % 2 ? -1 : 1;  /* boundary_normals, src/boundaries.c:11-18 */
        double nv = 0;
        for (int c = 0; c < 3; ++c) nv += nrm[c] * vels[c];
        for (int c = 0; c < 3; ++c) nrm[c] = nrm[c] * nv;
        for (int c = 0; c < 3; ++c) vels[c] = vels[c] - nrm[c];
    }
}

void ovo_velocity_direct_one(int n, const double *vnodes, const int *conn,
                             const int *status, const ovg_params *p, double t,
                             const double *tangents, const double *normals,
                             int k, double vs[3], double vels[3])
{
    double cnt = 0;
    velocity_one(n, vnodes, conn, status, p, t, tangents, normals, NULL, k,
                 vs, vels, &cnt);
}

void ovo_bs_direct_one(int n, const double *vnodes, const int *conn,
                       const ovg_params *p, int k, double out[3])
{
    const double *x = vnodes + 3 * k;
    out[0] = out[1] = out[2] = 0;
    for (int m = 0; m < n; ++m) {
        if (conn[2 * m] == -1 || k == m || k == conn[2 * m]) continue;
        double sv[3];
        segment_field(vnodes, conn, p, m, x, sv);
        for (int c = 0; c < 3; ++c) out[c] = out[c] + sv[c];
    }
    for (int j = 0; j < p->n_images; ++j) {
        double rs[3];
        shifted(x, j, p, rs);
        for (int m = 0; m < n; ++m) {
            if (conn[2 * m] == -1) continue;
            double sv[3];
            segment_field(vnodes, conn, p, m, rs, sv);
            for (int c = 0; c < 3; ++c) out[c] = out[c] + sv[c];
        }
    }
}

/* update_velocities, src/tangle.c:733-749 */
void ovo_velocities(int n, const double *vnodes, const int *conn,
                    const int *status, const ovg_params *p, double t,
                    const double *tangents, const double *normals,
                    double *vs, double *vels, int nthreads,
                    double *interactions)
{
    rtree tree, *tp = NULL;
    if (p->method == OVG_METHOD_BH) {
        rtree_build(&tree, n, vnodes, conn, status, p);
        tp = &tree;
    }
    double total = 0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16) reduction(+ : total)
#endif
    for (int k = 0; k < n; ++k) {
        double cnt = 0;
        velocity_one(n, vnodes, conn, status, p, t, tangents, normals, tp, k,
                     vs + 3 * k, vels + 3 * k, &cnt);
        total += cnt;
    }
    (void)nthreads;
    if (tp) rtree_free(tp);
    if (interactions) *interactions = total;
}

/* rk4_step, src/vortex_dynamics.c:50-119 with euler_step2 :20-42. */
void ovo_rk4_step(int n, double *vnodes, const int *conn, const int *status,
                  const ovg_params *p, double t, double dt,
                  const double *vels, int nthreads)
{
    size_t sz = sizeof(double) * 3 * (size_t)n;
    double *xs = (double *)calloc(1, sz), *tn = (double *)calloc(1, sz);
    double *nm = (double *)calloc(1, sz), *vsb = (double *)calloc(1, sz);
    double *k2 = (double *)calloc(1, sz), *k3 = (double *)calloc(1, sz);
    double *k4 = (double *)calloc(1, sz);
    const double *kin[3] = {vels, k2, k3};
    double *kout[3] = {k2, k3, k4};
    double tt[3] = {t + dt / 2, t + dt / 2, t + dt};
    double h4 = p->rk4_classic ? dt : dt / 2;
    double hs[3] = {dt / 2, dt / 2, h4};
    for (int s = 0; s < 3; ++s) {
        for (int k = 0; k < n; ++k) {
            if (status[2 * k] == OVG_EMPTY) continue;
            for (int c = 0; c < 3; ++c) {
                double move = kin[s][3 * k + c] * hs[s];
                xs[3 * k + c] = vnodes[3 * k + c] + move;
            }
        }
        ovo_tangents_normals(n, xs, conn, status, p, tn, nm);
        ovo_velocities(n, xs, conn, status, p, tt[s], tn, nm, vsb, kout[s],
                       nthreads, NULL);
    }
    for (int k = 0; k < n; ++k) {
        if (status[2 * k] == OVG_EMPTY) continue;
        for (int c = 0; c < 3; ++c) {
            double move = vels[3 * k + c] * (dt / 6);
            double a = k2[3 * k + c] * (dt / 3);
            move = move + a;
            a = k3[3 * k + c] * (dt / 3);
            move = move + a;
            a = k4[3 * k + c] * (dt / 6);
            move = move + a;
            vnodes[3 * k + c] = vnodes[3 * k + c] + move;
        }
    }
    free(xs); free(tn); free(nm); free(vsb); free(k2); free(k3); free(k4);
}

/* out_of_box, src/tangle.c:761-787 */
static int out_of_box(const double *x, const ovg_params *p)
{
    double b[6] = {p->box_lo[0], p->box_hi[0], p->box_lo[1],
                   p->box_hi[1], p->box_lo[2], p->box_hi[2]};
    int face = -1;
    if (x[0] < b[LEFT]) face = LEFT;
    if (x[0] > b[RIGHT]) face = RIGHT;
    if (x[1] < b[BACK]) face = BACK;
    if (x[1] > b[FRONT]) face = FRONT;
    if (x[2] < b[DOWN]) face = DOWN;
    if (x[2] > b[UP]) face = UP;
    if (face > -1 && p->wall[face] != OVG_WALL_PERIODIC) face = -1;
    return face;
}

/* enforce_periodic_boundaries, src/tangle.c:793-814 with periodic_shift,
 * src/boundaries.c:113-150 */
void ovo_enforce_periodic(int n, double *vnodes, const int *status,
                          const ovg_params *p)
{
    for (int k = 0; k < n; ++k) {
        if (status[2 * k] == OVG_EMPTY) continue;
        double *x = vnodes + 3 * k;
        int face;
        while ((face = out_of_box(x, p)) >= 0) {
            int c = face / 2;
            double L = p->box_hi[c] - p->box_lo[c];
            if (face % 2 == 0) x[c] += L;
            else x[c] -= L;
        }
    }
}

static int recon_eligible(const int *status, int k)
{
    int s = status[2 * k];
    return !(s == OVG_EMPTY || s == OVG_PINNED || s == OVG_PINNED_SLIP);
}

/* the predicate of src/vortex_dynamics.c:201-222 with all flags clear */
long ovo_reconnect_candidates(int n, const double *vnodes, const int *conn,
                              const int *status, const ovg_params *p,
                              long max_pairs, int *pairs)
{
    long cnt = 0;
    for (int k = 0; k < n; ++k) {
        if (!recon_eligible(status, k)) continue;
        for (int l = k + 1; l < n; ++l) {
            if (!recon_eligible(status, l) || conn[2 * k] == l ||
                conn[2 * k + 1] == l)
                continue;
            double w[3];
            pwrap(vnodes + 3 * k, vnodes + 3 * l, p, w);
            if (dist3(vnodes + 3 * k, w) > p->rec_dist) continue;
            if (cnt < max_pairs) {
                pairs[2 * cnt] = k;
                pairs[2 * cnt + 1] = l;
            }
            cnt++;
        }
    }
    return cnt;
}

/* do_reconnection, src/vortex_dynamics.c:128-175 */
static int do_reconnection(const double *vnodes, int *conn,
                           const ovg_params *p, int k, int l)
{
    int kf = conn[2 * k], kr = conn[2 * k + 1];
    int lf = conn[2 * l], lr = conn[2 * l + 1];
    double w[3];
#define PLEN(a, b) (pwrap(vnodes + 3 * (a), vnodes + 3 * (b), p, w), \
                    dist3(vnodes + 3 * (a), w))
    double l_kl = PLEN(k, l), l_kflr = PLEN(kf, lr), l_kkf = PLEN(k, kf);
    double l_llr = PLEN(l, lr), l_krlf = PLEN(kr, lf), l_kkr = PLEN(k, kr);
    double l_llf = PLEN(l, lf);
#undef PLEN
    double cf1 = l_kl + l_kflr - l_kkf - l_llr;
    double cf2 = l_kl + l_krlf - l_kkr - l_llf;
    if (cf1 > 0 && cf2 > 0) return 0;
    if (cf1 < cf2) {
        conn[2 * k] = l;
        conn[2 * l + 1] = k;
        conn[2 * kf + 1] = lr;
        conn[2 * lr] = kf;
    } else {
        conn[2 * l] = k;
        conn[2 * k + 1] = l;
        conn[2 * lf + 1] = kr;
        conn[2 * kr] = lf;
    }
    return 1;
}

/* reconnect, src/vortex_dynamics.c:186-288 */
int ovo_reconnect(int n, const double *vnodes, int *conn, const int *status,
                  const ovg_params *p, double t, int nthreads)
{
    (void)nthreads;
    long ncand = ovo_reconnect_candidates(n, vnodes, conn, status, p, 0, NULL);
    int *cand = (int *)malloc(sizeof(int) * 2 * (size_t)(ncand > 0 ? ncand : 1));
    ovo_reconnect_candidates(n, vnodes, conn, status, p, ncand, cand);
    int *flag = (int *)calloc((size_t)n, sizeof(int));
    size_t sz = sizeof(double) * 3 * (size_t)n;
    double *tn = (double *)calloc(1, sz), *nm = (double *)calloc(1, sz);
    int nrec = 0;
    double cutoff = cos(p->rec_angle);
    long i = 0;
    while (i < ncand) {
        int k = cand[2 * i];
        long iend = i;
        while (iend < ncand && cand[2 * iend] == k) ++iend;
        if (flag[k]) { i = iend; continue; }
        for (long j = i; j < iend; ++j) {
            int l = cand[2 * j + 1];
            if (flag[l]) continue;
            double w[3];
            pwrap(vnodes + 3 * k, vnodes + 3 * l, p, w);
            tangent_normal(vnodes, conn, status, p, k, tn + 3 * k, nm + 3 * k);
            tangent_normal(vnodes, conn, status, p, l, tn + 3 * l, nm + 3 * l);
            const double *d1 = tn + 3 * k, *d2 = tn + 3 * l;
            if (len3(d1) < 0.5 || len3(d2) < 0.5) continue;
            double calpha = ndot3(d1, d2);
            if (calpha > cutoff) continue;
            double vsk[3], vk[3], vsl[3], vl[3], cnt = 0;
            velocity_one(n, vnodes, conn, status, p, t, tn, nm, NULL, k, vsk,
                         vk, &cnt);
            velocity_one(n, vnodes, conn, status, p, t, tn, nm, NULL, l, vsl,
                         vl, &cnt);
            double dx[3], dv[3];
            for (int c = 0; c < 3; ++c) dx[c] = vnodes[3 * k + c] - w[c];
            for (int c = 0; c < 3; ++c) dv[c] = vk[c] - vl[c];
            if (dot3(dx, dv) > 0) continue;
            int knext = conn[2 * k], kprev = conn[2 * k + 1];
            int lnext = conn[2 * l], lprev = conn[2 * l + 1];
            if (!do_reconnection(vnodes, conn, p, k, l)) continue;
            flag[k]++; flag[conn[2 * k]]++; flag[conn[2 * k + 1]]++;
            flag[knext]++; flag[kprev]++;
            flag[l]++; flag[conn[2 * l]]++; flag[conn[2 * l + 1]]++;
            flag[lnext]++; flag[lprev]++;
            nrec++;
            break;
        }
        i = iend;
    }
    free(cand); free(flag); free(tn); free(nm);
    return nrec;
}

double ovo_line_length(int n, const double *vnodes, const int *conn,
                       const int *status, const ovg_params *p)
{
    double L = 0;
    for (int k = 0; k < n; ++k) {
        if (status[2 * k] == OVG_EMPTY || conn[2 * k] < 0) continue;
        double w[3];
        pwrap(vnodes + 3 * k, vnodes + 3 * conn[2 * k], p, w);
        L += dist3(vnodes + 3 * k, w);
    }
    return L;
}

/* octree_get_vs at arbitrary points with one reference-tree build. */
void ovo_bh_field(int n, const double *vnodes, const int *conn,
                  const int *status, const ovg_params *p, int m,
                  const double *pts, double *out)
{
    rtree tree;
    rtree_build(&tree, n, vnodes, conn, status, p);
    for (int i = 0; i < m; ++i) {
        double cnt = 0;
        rtree_vs(&tree, 0, pts + 3 * i, p->theta, p->kappa, out + 3 * i, &cnt);
    }
    rtree_free(&tree);
}

/* eliminate_small_loops, src/tangle.c:1063-1128, closed loops: a loop is
 * walked from its first unvisited slot counting m-1 steps for m points, and
 * deleted when that count is below the cutoff.  Returns deleted points. */
int ovo_eliminate_small_loops(int n, int *conn, int *status, int cutoff)
{
    char *visited = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    int removed = 0;
    for (int k = 0; k < n; ++k) {
        if (status[2 * k] == OVG_EMPTY || visited[k]) continue;
        visited[k] = 1;
        int loop = 0, here = k, next = conn[2 * k];
        while (next != k) {
            int sh = status[2 * here];
            if ((sh == OVG_PINNED || sh == OVG_PINNED_SLIP) && next < 0) {
                /* an open line: count back from k to its other end, on top
                 * of the forward steps (src/tangle.c:1081-1093) */
                here = k;
                next = conn[2 * here + 1];
                while (next >= 0) {
                    visited[here] = 1;
                    here = next;
                    next = conn[2 * here + 1];
                    loop++;
                }
                break;
            }
            visited[here] = 1;
            here = next;
            next = conn[2 * here];
            loop++;
        }
        if (loop < cutoff) {
            next = here;
            for (;;) {
                int tmp = next;
                next = conn[2 * next];
                conn[2 * tmp] = -1;
                conn[2 * tmp + 1] = -1;
                status[2 * tmp] = OVG_EMPTY;
                removed++;
                if (next == here || next < 0) break;
            }
        }
    }
    free(visited);
    return removed;
}

/* ---------------- remesh (src/tangle.c:1018-1061) ----------------------- */
/* search_next_empty_node, src/tangle.c:201-234 (a read past the last slot
 * counts as not EMPTY) */
static int search_next_empty(int N, const int *status, int *next_empty)
{
    if (*next_empty > 0) {
        if (*next_empty < N && status[2 * *next_empty] == OVG_EMPTY) {
            int r = *next_empty;
            if (*next_empty + 1 < N && status[2 * (*next_empty + 1)] == OVG_EMPTY)
                ++*next_empty;
            else
                *next_empty = -1;
            return r;
        }
    }
    for (int k = 0; k < N; ++k) {
        if (status[2 * k] == OVG_EMPTY) {
            *next_empty = (k + 1 < N && status[2 * (k + 1)] == OVG_EMPTY) ? k + 1 : -1;
            return k;
        }
    }
    return -1;
}

/* get_next_empty_node :241-253 with expand_tangle :68-101 (doubling) */
static int next_empty_node(int *N, int cap, int *conn, int *status, int *next_empty)
{
    int idx = search_next_empty(*N, status, next_empty);
    while (idx < 0 || status[2 * idx] != OVG_EMPTY) {
        int old = *N, nn = 2 * old;
        if (nn > cap) return -2;
        for (int k = old; k < nn; ++k) {
            status[2 * k] = OVG_EMPTY;
            status[2 * k + 1] = -1;
            conn[2 * k] = conn[2 * k + 1] = -1;
        }
        *N = nn;
        *next_empty = search_next_empty(*N, status, next_empty);
        idx = search_next_empty(*N, status, next_empty);
    }
    return idx;
}

/* add_point, src/tangle.c:261-335 */
static int add_point(int *N, int cap, double *vnodes, int *conn, int *status,
                     int *next_empty, const ovg_params *p, int pt)
{
    int nw = next_empty_node(N, cap, conn, status, next_empty);
    if (nw < 0) return nw;
    status[2 * nw] = OVG_FREE;
    status[2 * nw + 1] = -1;
    int next = conn[2 * pt];
    double t0[3], n0[3], t1[3], n1[3];
    tangent_normal(vnodes, conn, status, p, pt, t0, n0);
    tangent_normal(vnodes, conn, status, p, next, t1, n1);
    double s0[3], s1[3];
    memcpy(s0, vnodes + 3 * pt, 24);
    pwrap(s0, vnodes + 3 * next, p, s1);
    double l = dist3(s0, s1);
    double nv[3], nwp[3];
    for (int c = 0; c < 3; ++c) nv[c] = n0[c] + n1[c];
    for (int c = 0; c < 3; ++c) nv[c] = nv[c] * 0.5;
    if (len3(nv) > 1e-5) {
        double R = 1 / len3(nv);
        double dR = R * R - l * l / 4;
        double delta = dR > 0 ? R - sqrt(dR) : R;
        double ln = len3(nv);
        for (int c = 0; c < 3; ++c) nv[c] = nv[c] * (1 / ln);
        for (int c = 0; c < 3; ++c) nv[c] = nv[c] * -1;
        double a[3], b[3];
        for (int c = 0; c < 3; ++c) a[c] = s0[c] + s1[c];
        for (int c = 0; c < 3; ++c) a[c] = a[c] * 0.5;
        for (int c = 0; c < 3; ++c) b[c] = nv[c] * delta;
        for (int c = 0; c < 3; ++c) nwp[c] = a[c] + b[c];
    } else {
        for (int c = 0; c < 3; ++c) nwp[c] = s0[c] + s1[c];
        for (int c = 0; c < 3; ++c) nwp[c] = nwp[c] * 0.5;
    }
    memcpy(vnodes + 3 * nw, nwp, 24);
    conn[2 * nw + 1] = pt;
    conn[2 * nw] = next;
    conn[2 * pt] = nw;
    conn[2 * next + 1] = nw;
    return nw;
}

/* remesh; *N is the reference's tangle->N (may double up to cap), arrays
 * must hold cap slots.  Returns points added, -1 on capacity overflow;
 * *removed receives the removed count. */
int ovo_remesh(int *N, int cap, double *vnodes, int *conn, int *status,
               int *next_empty, const ovg_params *p, double min_dist,
               double max_dist, int *removed)
{
    int added = 0, rem = 0;
    for (int k = 0; k < *N; ++k) {
        if (status[2 * k] == OVG_EMPTY) continue;
        int next = conn[2 * k], prev = conn[2 * k + 1];
        double lf = 0, lr = 0, w[3];
        if (next >= 0) {
            pwrap(vnodes + 3 * k, vnodes + 3 * next, p, w);
            lf = dist3(vnodes + 3 * k, w);
        }
        if (prev >= 0) {
            pwrap(vnodes + 3 * k, vnodes + 3 * prev, p, w);
            lr = dist3(vnodes + 3 * k, w);
        }
        if (next >= 0 && prev >= 0 && ((lf < min_dist || lr < min_dist) && (lf + lr) < max_dist)) {
            /* remove_point, src/tangle.c:342-355 */
            int pv = conn[2 * k + 1], nx = conn[2 * k];
            if (pv >= 0) conn[2 * pv] = nx;
            if (nx >= 0) conn[2 * nx + 1] = pv;
            conn[2 * k] = conn[2 * k + 1] = -1;
            status[2 * k] = OVG_EMPTY;
            rem++;
        }
        if (next >= 0 && lf > max_dist) {
            added++;
            if (add_point(N, cap, vnodes, conn, status, next_empty, p, k) < 0) return -1;
        }
    }
    if (added) ovo_enforce_periodic(*N, vnodes, status, p);
    if (removed) *removed = rem;
    return added;
}
