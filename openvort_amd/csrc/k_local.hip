// Per-point streaming kernels: the 5-point tangent/normal stencil, the
// direct-sum source preparation, the velocity composition (LIA + external
// flow + Biot-Savart + mutual friction) fused with the RK4 stage updates,
// the periodic wrap and the line length.  All FP64 in the reference's
// operation order; HBM-bound.
#include "device_local.h"

namespace ovg {

__global__ void k_stencil(int n_live, const int *__restrict__ live, const double *__restrict__ X,
                          const double *__restrict__ Y, const double *__restrict__ Z,
                          const int *__restrict__ fwd, const int *__restrict__ rev,
                          const int *__restrict__ st, const int *__restrict__ pinw,
                          const TangleParams *__restrict__ Pp, double *__restrict__ tx,
                          double *__restrict__ ty, double *__restrict__ tz,
                          double *__restrict__ nx, double *__restrict__ ny,
                          double *__restrict__ nz)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_live) return;
    int k = live[i];
    const TangleParams &P = tparams_of(Pp, k);
    double t[3], n[3];
    tangent_normal(P, X, Y, Z, fwd, rev, st, pinw, k, t, n);
    tx[k] = t[0]; ty[k] = t[1]; tz[k] = t[2];
    nx[k] = n[0]; ny[k] = n[1]; nz[k] = n[2];
}

void launch_stencil(Ctx &c, const double *const x[3], double *const tan[3],
                    double *const nor[3])
{
    if (c.n_live == 0) return;
    int bs = 256, gs = (c.n_live + bs - 1) / bs;
    hipLaunchKernelGGL(k_stencil, dim3(gs), dim3(bs), 0, c.stream, c.n_live,
                       c.live.as<int>(), x[0], x[1], x[2], c.fwd.as<int>(), c.rev.as<int>(),
                       c.status.as<int>(), c.has_pinned ? c.pinw.as<int>() : nullptr,
                       dparams(c), tan[0], tan[1], tan[2], nor[0], nor[1], nor[2]);
}

// Direct-sum sources: r1 = x_m and the pwrapped r2 = seg_pwrap(x_m,
// x_fwd(m)).r2, exactly the segment built inside segment_field
// (src/tangle.c:534), once per evaluation instead of once per pair.
__global__ void k_prep_segments(int n_src, const int *__restrict__ src,
                                const double *__restrict__ X, const double *__restrict__ Y,
                                const double *__restrict__ Z, const int *__restrict__ fwd,
                                const TangleParams *__restrict__ Pp, double *__restrict__ s1x,
                                double *__restrict__ s1y, double *__restrict__ s1z,
                                double *__restrict__ s2x, double *__restrict__ s2y,
                                double *__restrict__ s2z, int2 *__restrict__ ids)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_src) return;
    const TangleParams &P = *Pp;
    int m = src[j];
    int f = fwd[m];
    double ax = X[m], ay = Y[m], az = Z[m];
    // an open end (a pinned node with no forward link) has no segment
    // (src/tangle.c:658): a zero-length one adds exactly zero
    double bx = ax, by = ay, bz = az;
    if (f >= 0) {
        bx = X[f]; by = Y[f]; bz = Z[f];
        pwrap(P, ax, ay, az, bx, by, bz);
    }
    s1x[j] = ax; s1y[j] = ay; s1z[j] = az;
    s2x[j] = bx; s2y[j] = by; s2z[j] = bz;
    ids[j] = make_int2(m, f);
}

void launch_prep_segments(Ctx &c, const double *const x[3])
{
    if (c.n_src == 0) return;
    int bs = 256, gs = (c.n_src + bs - 1) / bs;
    hipLaunchKernelGGL(k_prep_segments, dim3(gs), dim3(bs), 0, c.stream, c.n_src,
                       c.src.as<int>(), x[0], x[1], x[2], c.fwd.as<int>(), dparams(c),
                       c.seg[0].as<double>(), c.seg[1].as<double>(), c.seg[2].as<double>(),
                       c.seg[3].as<double>(), c.seg[4].as<double>(), c.seg[5].as<double>(),
                       c.seg_ids.as<int2>());
}

struct ComposeArgs {
    int n_live;
    const int *live;
    const double *X, *Y, *Z;          // evaluation positions
    const int *fwd, *rev;
    const int *st, *pinw;             // status, pin walls (pinw null: no pinned nodes)
    const double *tin[3], *nin[3];    // tangents/normals when not computed here
    double *tout[3], *nout[3];        // where computed tangents go (may be null)
    const double *bsp[3], *bsi[3];    // Biot-Savart: primary and images
    const TangleParams *P;
    double *vs[3];                    // may be null
    double *vels[3];                  // may be null
    // RK4 fusion: stage 0 = plain evaluation, 1..3 = rk4 stages 2..4
    int stage;
    double h_next;                    // Euler step for the next stage position
    double w_stage;                   // this stage's combine weight (dt/3 or dt/6)
    double w_k1;                      // dt/6, stage 1 only
    const double *k1[3];              // tangle vels (k1)
    double *x0[3];                    // base positions (updated at stage 3)
    double *move[3];
    double *xnext[3];
};

// update_velocity composition, src/tangle.c:600-726 (FREE nodes):
//   vs   = LIA + v_s,ext + BS_primary + BS_images
//   vels = vs + alpha s' x (v_n - vs) - alpha' s' x (s' x (v_n - vs))
// with lia_velocity src/tangle.c:568-592; then, for RK4 stages, the
// combine of src/vortex_dynamics.c:95-115 and the next euler_step2
// (:20-42), accumulated in the reference's order.
template <bool OWN_TAN>
__global__ void k_compose(ComposeArgs a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_live) return;
    int k = a.live[i];
    if (k < 0) return;  // a target slot the caller's list had wrong (checked copy)
    const TangleParams &P = tparams_of(a.P, k);
    double t[3], nr[3];
    if (OWN_TAN) {
        tangent_normal(P, a.X, a.Y, a.Z, a.fwd, a.rev, a.st, a.pinw, k, t, nr);
        if (a.tout[0]) {
            for (int c = 0; c < 3; ++c) {
                a.tout[c][k] = t[c];
                a.nout[c][k] = nr[c];
            }
        }
    } else {
        for (int c = 0; c < 3; ++c) {
            t[c] = a.tin[c][k];
            nr[c] = a.nin[c][k];
        }
    }
    // lia_velocity
    double px = a.X[k], py = a.Y[k], pz = a.Z[k];
    double fx, fy, fz, qx, qy, qz;
    const int sk = a.pinw ? a.st[k] : OVG_FREE;
    if (a.pinw) {
        step_pos(P, a.X, a.Y, a.Z, a.fwd, a.rev, a.st, a.pinw, k, 1, fx, fy, fz);
        step_pos(P, a.X, a.Y, a.Z, a.fwd, a.rev, a.st, a.pinw, k, -1, qx, qy, qz);
    } else {
        int inx = a.fwd[k], ipr = a.rev[k];
        fx = a.X[inx]; fy = a.Y[inx]; fz = a.Z[inx];
        qx = a.X[ipr]; qy = a.Y[ipr]; qz = a.Z[ipr];
    }
    pwrap(P, px, py, pz, fx, fy, fz);           // seg_pwrap(p, next)
    double wx = px, wy = py, wz = pz;
    pwrap(P, qx, qy, qz, wx, wy, wz);           // seg_pwrap(prev, p)
    double l_next = dist3(px, py, pz, fx, fy, fz);
    double l_prev = dist3(qx, qy, qz, wx, wy, wz);
    double f = P.kappa * log(2 * sqrt(l_next * l_prev) / P.sqrtE / P.a0) / 4 / M_PI;
    double cx = t[1] * nr[2] - t[2] * nr[1];
    double cy = t[2] * nr[0] - t[0] * nr[2];
    double cz = t[0] * nr[1] - t[1] * nr[0];
    double v[3] = {cx * f, cy * f, cz * f};
    for (int c = 0; c < 3; ++c) v[c] = v[c] + P.evs[c];
    for (int c = 0; c < 3; ++c) v[c] = v[c] + a.bsp[c][k];
    for (int c = 0; c < 3; ++c) v[c] = v[c] + a.bsi[c][k];
    double ve[3] = {v[0], v[1], v[2]};
    if (P.use_mf) {
        double dv[3];
        for (int c = 0; c < 3; ++c) dv[c] = P.evn[c] - v[c];
        double m0 = t[1] * dv[2] - t[2] * dv[1];
        double m1 = t[2] * dv[0] - t[0] * dv[2];
        double m2 = t[0] * dv[1] - t[1] * dv[0];
        ve[0] = ve[0] + m0 * P.alpha;
        ve[1] = ve[1] + m1 * P.alpha;
        ve[2] = ve[2] + m2 * P.alpha;
        double n0 = t[1] * m2 - t[2] * m1;
        double n1 = t[2] * m0 - t[0] * m2;
        double n2 = t[0] * m1 - t[1] * m0;
        double ap = -P.alpha_p;
        ve[0] = ve[0] + n0 * ap;
        ve[1] = ve[1] + n1 * ap;
        ve[2] = ve[2] + n2 * ap;
    }
    if (sk == OVG_PINNED) {
        // src/tangle.c:608-615: no self-induced motion, the boundary's velocity
        for (int c = 0; c < 3; ++c) {
            v[c] = 0;
            ve[c] = P.evb[c];
        }
    } else if (sk == OVG_PINNED_SLIP) {
        // src/tangle.c:718-725: the velocity through the wall removed;
        // boundary_normals (src/boundaries.c:11-18) are +-e_axis
        const int w = a.pinw[k], ax = w >> 1;
        double n3[3] = {0, 0, 0};
        n3[ax] = (w & 1) ? -1.0 : 1.0;
        double nv = dot3(n3[0], n3[1], n3[2], ve[0], ve[1], ve[2]);
        for (int c = 0; c < 3; ++c) ve[c] = ve[c] - n3[c] * nv;
    }
    if (a.vs[0])
        for (int c = 0; c < 3; ++c) a.vs[c][k] = v[c];
    if (a.vels[0])
        for (int c = 0; c < 3; ++c) a.vels[c][k] = ve[c];
    if (a.stage == 1) {
        for (int c = 0; c < 3; ++c) {
            double mv = a.k1[c][k] * a.w_k1;
            double ad = ve[c] * a.w_stage;
            a.move[c][k] = mv + ad;
            a.xnext[c][k] = a.x0[c][k] + ve[c] * a.h_next;
        }
    } else if (a.stage == 2) {
        for (int c = 0; c < 3; ++c) {
            a.move[c][k] = a.move[c][k] + ve[c] * a.w_stage;
            a.xnext[c][k] = a.x0[c][k] + ve[c] * a.h_next;
        }
    } else if (a.stage == 3) {
        for (int c = 0; c < 3; ++c) {
            double mv = a.move[c][k] + ve[c] * a.w_stage;
            a.x0[c][k] = a.x0[c][k] + mv;
        }
    }
}

// stage: 0 plain evaluation at c.x writing c.vs/c.vels (and tangents when
// own_tangents), 1..3 RK4 stages at the positions x.
void launch_compose(Ctx &c, const double *const x[3], bool own_tangents, int stage, double dt)
{
    if (c.n_live == 0) return;
    ComposeArgs a{};
    a.n_live = c.n_live;
    a.live = c.live.as<int>();
    a.X = x[0]; a.Y = x[1]; a.Z = x[2];
    a.fwd = c.fwd.as<int>();
    a.rev = c.rev.as<int>();
    a.st = c.status.as<int>();
    a.pinw = c.has_pinned ? c.pinw.as<int>() : nullptr;
    a.P = dparams(c);
    for (int q = 0; q < 3; ++q) {
        a.tin[q] = c.tan[q].as<double>();
        a.nin[q] = c.nor[q].as<double>();
        a.bsp[q] = c.bs_prim[q].as<double>();
        a.bsi[q] = c.bs_img[q].as<double>();
        a.k1[q] = c.vels[q].as<double>();
        a.x0[q] = c.x[q].as<double>();
        a.move[q] = c.move[q].as<double>();
    }
    a.stage = stage;
    if (stage == 0) {
        for (int q = 0; q < 3; ++q) {
            a.vs[q] = c.vs[q].as<double>();
            a.vels[q] = c.vels[q].as<double>();
            a.tout[q] = own_tangents ? c.tan[q].as<double>() : nullptr;
            a.nout[q] = own_tangents ? c.nor[q].as<double>() : nullptr;
        }
    } else {
        double h4 = c.params.rk4_classic ? dt : dt / 2;
        a.w_k1 = dt / 6;
        a.w_stage = stage == 3 ? dt / 6 : dt / 3;
        a.h_next = stage == 1 ? dt / 2 : h4;
        int nb = stage == 1 ? 1 : 0;  // stage 1 writes xs[1], stage 2 writes xs[0]
        for (int q = 0; q < 3; ++q) {
            a.vs[q] = nullptr;
            a.vels[q] = nullptr;
            a.tout[q] = nullptr;
            a.nout[q] = nullptr;
            a.xnext[q] = c.xs[nb][q].as<double>();
        }
    }
    int bs = 256, gs = (c.n_live + bs - 1) / bs;
    if (own_tangents)
        hipLaunchKernelGGL(k_compose<true>, dim3(gs), dim3(bs), 0, c.stream, a);
    else
        hipLaunchKernelGGL(k_compose<false>, dim3(gs), dim3(bs), 0, c.stream, a);
}

// euler_step2, src/vortex_dynamics.c:20-42: out = x + v * h
__global__ void k_stage_pos(int n_live, const int *__restrict__ live,
                            const double *__restrict__ X, const double *__restrict__ Y,
                            const double *__restrict__ Z, const double *__restrict__ vx,
                            const double *__restrict__ vy, const double *__restrict__ vz,
                            double h, double *__restrict__ ox, double *__restrict__ oy,
                            double *__restrict__ oz)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_live) return;
    int k = live[i];
    ox[k] = X[k] + vx[k] * h;
    oy[k] = Y[k] + vy[k] * h;
    oz[k] = Z[k] + vz[k] * h;
}

void launch_stage_positions(Ctx &c, const double *const kv[3], double h,
                            double *const out[3])
{
    if (c.n_live == 0) return;
    int bs = 256, gs = (c.n_live + bs - 1) / bs;
    hipLaunchKernelGGL(k_stage_pos, dim3(gs), dim3(bs), 0, c.stream, c.n_live,
                       c.live.as<int>(), c.x[0].as<double>(), c.x[1].as<double>(),
                       c.x[2].as<double>(), kv[0], kv[1], kv[2], h, out[0], out[1], out[2]);
}

// enforce_periodic_boundaries, src/tangle.c:793-814 with out_of_box
// :761-787 and periodic_shift src/boundaries.c:113-150
// live == nullptr: every slot that is not EMPTY (st), and only when the
// device count *gate is nonzero.
__global__ void k_wrap(int n_live, const int *__restrict__ live, double *__restrict__ X,
                       double *__restrict__ Y, double *__restrict__ Z,
                       const TangleParams *__restrict__ Pp, int *__restrict__ flags,
                       const int *__restrict__ st, const int *__restrict__ gate)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_live) return;
    if (!live && (*gate == 0 || st[i] == OVG_EMPTY)) return;
    int k = live ? live[i] : i;
    const TangleParams &P = tparams_of(Pp, k);
    double p[3] = {X[k], Y[k], Z[k]};
    int moved = 0;
    for (int it = 0; it < 256; ++it) {
        int face = -1;
        if (p[0] < P.lo[0]) face = 0;
        if (p[0] > P.hi[0]) face = 1;
        if (p[1] < P.lo[1]) face = 2;
        if (p[1] > P.hi[1]) face = 3;
        if (p[2] < P.lo[2]) face = 4;
        if (p[2] > P.hi[2]) face = 5;
        if (face > -1 && !P.wrap_face[face]) face = -1;
        if (face < 0) break;
        int cc = face >> 1;
        if ((face & 1) == 0) p[cc] += P.L[cc];
        else p[cc] -= P.L[cc];
        moved = 1;
        if (it == 255) atomicOr(flags, 1);
    }
    X[k] = p[0]; Y[k] = p[1]; Z[k] = p[2];
    // flags[1]: points moved by this wrap (one atomic per wavefront)
    const unsigned long long mv = __ballot(moved);
    if ((threadIdx.x & 63) == __ffsll((long long)__ballot(1)) - 1 && mv)
        atomicAdd(flags + 1, __popcll(mv));
}

__global__ void k_zero_int(int *a)
{
    if (threadIdx.x == 0) *a = 0;
}

void launch_wrap(Ctx &c)
{
    if (c.n_live == 0) return;
    int bs = 256, gs = (c.n_live + bs - 1) / bs;
    // a current host position mirror stays current unless this wrap moves
    // points (host_positions reads the count); a copy in flight finishes first
    if (c.xpin_pending) {
        hipEventSynchronize(c.xpin_done);
        c.xpin_pending = false;
        c.xpin_state = 1;
    }
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c.stream, c.flags.as<int>() + 1);
    if (c.xpin_state != 0) c.xpin_state = 2;
    hipLaunchKernelGGL(k_wrap, dim3(gs), dim3(bs), 0, c.stream, c.n_live, c.live.as<int>(),
                       c.x[0].as<double>(), c.x[1].as<double>(), c.x[2].as<double>(),
                       dparams(c), c.flags.as<int>(), nullptr, nullptr);
}

void launch_wrap_slots(Ctx &c, const int *d_gate)
{
    if (c.n == 0) return;
    int bs = 256, gs = (c.n + bs - 1) / bs;
    hipLaunchKernelGGL(k_wrap, dim3(gs), dim3(bs), 0, c.stream, c.n, nullptr,
                       c.x[0].as<double>(), c.x[1].as<double>(), c.x[2].as<double>(),
                       dparams(c), c.flags.as<int>(), c.status.as<int>(), d_gate);
}

// sum over live slots of |seg_pwrap(x_k, x_fwd(k))|: per-block partials,
// summed on the host in block order (deterministic).
__global__ void k_line_length(int n_live, const int *__restrict__ live,
                              const double *__restrict__ X, const double *__restrict__ Y,
                              const double *__restrict__ Z, const int *__restrict__ fwd,
                              const TangleParams *__restrict__ Pp, double *__restrict__ out)
{
    __shared__ double red[256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0;
    if (i < n_live) {
        int k = live[i];
        const TangleParams &P = tparams_of(Pp, k);
        int f = fwd[k];
        if (f >= 0) {
            double bx = X[f], by = Y[f], bz = Z[f];
            pwrap(P, X[k], Y[k], Z[k], bx, by, bz);
            l = dist3(X[k], Y[k], Z[k], bx, by, bz);
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

void launch_line_length(Ctx &c, double *d_out)
{
    int bs = 256, gs = (c.n_live + bs - 1) / bs;
    if (gs == 0) return;
    hipLaunchKernelGGL(k_line_length, dim3(gs), dim3(bs), 0, c.stream, c.n_live,
                       c.live.as<int>(), c.x[0].as<double>(), c.x[1].as<double>(),
                       c.x[2].as<double>(), c.fwd.as<int>(), dparams(c), d_out);
}

// Flags bit 1 when any live position is not finite.
__global__ void k_check_finite(int n_live, const int *__restrict__ live,
                               const double *__restrict__ X, const double *__restrict__ Y,
                               const double *__restrict__ Z, int *__restrict__ flags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_live) return;
    int k = live[i];
    if (!(isfinite(X[k]) && isfinite(Y[k]) && isfinite(Z[k]))) atomicOr(flags, 2);
}

void launch_check_finite(Ctx &c, const double *const x[3])
{
    if (c.n_live == 0) return;
    int bs = 256, gs = (c.n_live + bs - 1) / bs;
    hipLaunchKernelGGL(k_check_finite, dim3(gs), dim3(bs), 0, c.stream, c.n_live,
                       c.live.as<int>(), x[0], x[1], x[2], c.flags.as<int>());
}

}  // namespace ovg
