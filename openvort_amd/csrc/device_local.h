// Per-point FP64 stencil helpers shared by the streaming kernels and the
// device topology passes: link walks, step_node positions and the
// tangent/normal stencil, in the reference's operation order.
#pragma once
#include "engine_internal.h"

namespace ovg {

// walk `steps` links forward (>0) or backward (<0): step_node for FREE
// nodes, src/tangle.c:155-168
__device__ __forceinline__ int walk(const int *__restrict__ fwd, const int *__restrict__ rev,
                                    int i, int steps)
{
    for (; steps > 0; --steps) i = fwd[i];
    for (; steps < 0; ++steps) i = rev[i];
    return i;
}

__device__ __forceinline__ double dist3(double ax, double ay, double az, double bx, double by,
                                        double bz)
{
    double dx = ax - bx, dy = ay - by, dz = az - bz;
    return sqrt(dot3(dx, dy, dz, dx, dy, dz));
}

// step_node, src/tangle.c:155-194, as a position: `where` links from k.  A
// pinned node at its open end walks the other way instead and the result
// is mirrored through its wall (mirror_shift, src/boundaries.c:159-193);
// the mirrors of nested ends apply innermost first.  A pinned node with no
// link at all has no stencil: NaN.
static __device__ void step_pos(const TangleParams &P, const double *__restrict__ X,
                         const double *__restrict__ Y, const double *__restrict__ Z,
                         const int *__restrict__ fwd, const int *__restrict__ rev,
                         const int *__restrict__ st, const int *__restrict__ pinw, int k,
                         int where, double &x, double &y, double &z)
{
    int walls[4], nw = 0, i = k;
    for (int guard = 0; where != 0 && guard < 8; ++guard) {
        const int s = st[i];
        const int nx = where > 0 ? fwd[i] : rev[i];
        if ((s == OVG_PINNED || s == OVG_PINNED_SLIP) && nx < 0) {
            if (nw < 4) walls[nw++] = pinw[i];
            where = -where;
            continue;
        }
        i = nx;
        where += where > 0 ? -1 : 1;
    }
    x = X[i]; y = Y[i]; z = Z[i];
    if (where != 0) x = y = z = __builtin_nan("");
    for (int j = nw - 1; j >= 0; --j) {
        const int w = walls[j], a = w >> 1;
        const double wall = (w & 1) ? P.hi[a] : P.lo[a];
        if (a == 0) x = 2 * wall - x;
        else if (a == 1) y = 2 * wall - y;
        else z = 2 * wall - z;
    }
}

// update_tangent_normal's arithmetic, src/tangle.c:446-496, from the
// positions of the five stencil points s0 (the point), s1, s2 (one and two
// links forward) and m1, m2 (backward).
static __device__ void tangent_normal_pts(
    const TangleParams &P, double s0x, double s0y, double s0z, double s1x, double s1y,
    double s1z, double s2x, double s2y, double s2z, double m1x, double m1y, double m1z,
    double m2x, double m2y, double m2z, double t[3], double nrm[3])
{
    double w[4][3];
    w[0][0] = s2x; w[0][1] = s2y; w[0][2] = s2z;
    pwrap(P, s0x, s0y, s0z, w[0][0], w[0][1], w[0][2]);
    w[1][0] = s1x; w[1][1] = s1y; w[1][2] = s1z;
    pwrap(P, s0x, s0y, s0z, w[1][0], w[1][1], w[1][2]);
    w[2][0] = m1x; w[2][1] = m1y; w[2][2] = m1z;
    pwrap(P, s0x, s0y, s0z, w[2][0], w[2][1], w[2][2]);
    w[3][0] = m2x; w[3][1] = m2y; w[3][2] = m2z;
    pwrap(P, s0x, s0y, s0z, w[3][0], w[3][1], w[3][2]);
    double w12x = s2x, w12y = s2y, w12z = s2z;
    pwrap(P, s1x, s1y, s1z, w12x, w12y, w12z);
    double wm12x = m2x, wm12y = m2y, wm12z = m2z;
    pwrap(P, m1x, m1y, m1z, wm12x, wm12y, wm12z);
    double ds[4][3];
    for (int j = 0; j < 4; ++j) {
        ds[j][0] = w[j][0] - s0x;
        ds[j][1] = w[j][1] - s0y;
        ds[j][2] = w[j][2] - s0z;
    }
    double d1 = dist3(s0x, s0y, s0z, w[1][0], w[1][1], w[1][2]);
    double d2 = d1 + dist3(s1x, s1y, s1z, w12x, w12y, w12z);
    double dm1 = dist3(s0x, s0y, s0z, w[2][0], w[2][1], w[2][2]);
    double dm2 = dm1 + dist3(m1x, m1y, m1z, wm12x, wm12y, wm12z);
    double dsd[4] = {d2 * (d2 - d1) * (dm1 + d2) * (dm2 + d2),
                     d1 * (d2 - d1) * (dm1 + d1) * (dm2 + d1),
                     dm1 * (dm1 + d1) * (dm1 + d2) * (dm2 - dm1),
                     dm2 * (dm2 + d1) * (dm2 + d2) * (dm2 - dm1)};
    double c1[4] = {-d1 * dm1 * dm2, d2 * dm1 * dm2, -d1 * d2 * dm2, d1 * d2 * dm1};
    double c2[4] = {2 * ((dm1 - d1) * dm2 - d1 * dm1), -2 * ((dm1 - d2) * dm2 - d2 * dm1),
                    2 * ((d2 + d1) * dm2 - d1 * d2), -2 * ((d2 + d1) * dm1 - d1 * d2)};
    double a1[4], a2[4];
    for (int z = 0; z < 4; ++z) {
        a1[z] = c1[z] / dsd[z];
        a2[z] = c2[z] / dsd[z];
    }
    for (int i = 0; i < 3; ++i) {
        double tt = a1[0] * ds[0][i];
        double nn = a2[0] * ds[0][i];
        for (int z = 1; z < 4; ++z) {
            tt = tt + a1[z] * ds[z][i];
            nn = nn + a2[z] * ds[z][i];
        }
        t[i] = tt;
        nrm[i] = nn;
    }
}

// update_tangent_normal, src/tangle.c:423-502.  pinw == nullptr: no pinned
// nodes, every neighbour is a link away.
static __device__ void tangent_normal(const TangleParams &P, const double *__restrict__ X,
                               const double *__restrict__ Y, const double *__restrict__ Z,
                               const int *__restrict__ fwd, const int *__restrict__ rev,
                               const int *__restrict__ st, const int *__restrict__ pinw,
                               int k, double t[3], double nrm[3])
{
    double s0x = X[k], s0y = Y[k], s0z = Z[k];
    double s1x, s1y, s1z, s2x, s2y, s2z, m1x, m1y, m1z, m2x, m2y, m2z;
    if (pinw) {
        step_pos(P, X, Y, Z, fwd, rev, st, pinw, k, 1, s1x, s1y, s1z);
        step_pos(P, X, Y, Z, fwd, rev, st, pinw, k, 2, s2x, s2y, s2z);
        step_pos(P, X, Y, Z, fwd, rev, st, pinw, k, -1, m1x, m1y, m1z);
        step_pos(P, X, Y, Z, fwd, rev, st, pinw, k, -2, m2x, m2y, m2z);
    } else {
        int i1 = fwd[k], i2 = fwd[i1], im1 = rev[k], im2 = rev[im1];
        s1x = X[i1]; s1y = Y[i1]; s1z = Z[i1];
        s2x = X[i2]; s2y = Y[i2]; s2z = Z[i2];
        m1x = X[im1]; m1y = Y[im1]; m1z = Z[im1];
        m2x = X[im2]; m2y = Y[im2]; m2z = Z[im2];
    }
    tangent_normal_pts(P, s0x, s0y, s0z, s1x, s1y, s1z, s2x, s2y, s2z, m1x, m1y, m1z, m2x, m2y,
                       m2z, t, nrm);
}

}  // namespace ovg
