// Topology passes on the device (ovg_params.topology == OVG_TOPO_DEVICE):
//   reconnect              src/vortex_dynamics.c:186-288, do_reconnection :128-175
//   eliminate_small_loops  src/tangle.c:1063-1128 (closed loops)
//   remesh                 src/tangle.c:1018-1061 with add_point :261-335,
//                          remove_point :342-355 and get_next_empty_node :201-253
// as kernels on the links, status and positions in HBM.  Nothing here waits
// for the host: the passes leave their counts in a small array of device
// counters (DT_*) that the host reads when a call needs them (sync_live).
//
// The sequential parts run in one wavefront and keep the reference's
// visiting order.  They make the same decisions as the host passes
// (host_recon.h, engine.hip) with the same FP64 operation order, so the two
// give identical links, status and positions; what runs in parallel around
// them (candidate search, endpoint stencils and velocities, the remesh
// candidate test, the small-loop walks) only computes inputs that the
// sequential order cannot change.
//
// In the single-wavefront kernels every lane runs the same scalar code and
// stores the same values; each lane therefore reads back its own stores.
// Arrays those kernels write are read with relaxed atomic loads, which are
// never served from the scalar cache.
#include <algorithm>
#include <climits>
#include <cstring>

#include "device_local.h"

namespace ovg {

// device counters (int, c.dt_state)
enum {
    DT_LIVE = 0,   // live points after the device passes (read by sync_live)
    DT_ERR,        // error bits of the passes since the last sync_live
    DT_HINT,       // the allocator hint, tangle->next_empty
    DT_LO,         // remesh: every slot below is not EMPTY
    DT_REC,        // reconnections of the last reconnection pass
    DT_LOOPS,      // points removed by the last small-loop pass
    DT_ADDED,      // points added by the last remesh
    DT_REMOVED,    // points removed by the last remesh
    DT_PAIRS,      // candidate pairs the last reconnection pass took
    DT_FOUND,      // candidate pairs it found (more than DT_PAIRS: over capacity)
    DT_ENDS,       // their distinct endpoints
    DT_CHANGES,    // segments rewired by the last reconnection pass
    DT_GROUPS,     // endpoint target groups of the list traversal
    DT_COUNT = 16
};
constexpr int kErrPairs = 1, kErrEnds = 2, kErrSlots = 4, kErrHeap = 8;
constexpr int kMaxPairs = 32768, kMaxEnds = 65536;
constexpr int kNear = 8;          // tree mode: rewirable pairs listed per endpoint
constexpr int kHeapLds = 8192;    // remesh: top of the heap of touched slots, in LDS
constexpr int kGroupSort = 8192;  // endpoints grouped by tree order in one LDS sort

template <class T>
__device__ __forceinline__ T ld(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ double dt_dist(const TangleParams &P, double ax, double ay, double az,
                                          double bx, double by, double bz)
{
    // |seg_pwrap(a, b)|: b moved next to a, then the distance
    pwrap(P, ax, ay, az, bx, by, bz);
    return dist3(ax, ay, az, bx, by, bz);
}

__global__ void k_dt_set(int *__restrict__ a, int i, int v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) a[i] = v;
}

__global__ void k_dt_fill(int n, int *__restrict__ a, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// one atomic per wavefront for a per-lane count
__device__ __forceinline__ void wave_add(int *p, bool on)
{
    const unsigned long long b = __ballot(on);
    if (b && (threadIdx.x & 63) == __ffsll((long long)__ballot(1)) - 1)
        atomicAdd(p, (int)__popcll(b));
}

// ------------------------------------------------------------ small loops
// A closed loop of m points is erased exactly when m - 1 < cutoff: the
// reference walks the links from k back to k counting m - 1 of them
// (src/tangle.c:1079-1099) and erases when that count is below the cutoff
// (:1101-1118).  Loops are disjoint, so each slot decides for its own loop
// from the links before the pass, whatever order the reference visits.
__global__ void k_dt_loop_flag(int n, const int *__restrict__ fwd, const int *__restrict__ st,
                               int cutoff, int *__restrict__ flag)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int f = 0;
    if (st[k] != OVG_EMPTY) {
        int next = fwd[k], loop = 0;
        while (next != k && next >= 0 && loop < cutoff) {
            next = fwd[next];
            ++loop;
        }
        f = next == k && loop < cutoff;
    }
    flag[k] = f;
}

__global__ void k_dt_loop_erase(int n, const int *__restrict__ flag, int *__restrict__ fwd,
                                int *__restrict__ rev, int *__restrict__ st,
                                int *__restrict__ state)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = k < n && flag[k];
    if (on) {
        fwd[k] = -1;
        rev[k] = -1;
        st[k] = OVG_EMPTY;
    }
    wave_add(state + DT_LOOPS, on);
}

// ----------------------------------------------------------------- remesh
// lowest EMPTY slot (n when none): the start of get_next_empty_node's scan
__global__ void k_dt_first_empty(int n, const int *__restrict__ st, int *__restrict__ state)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long b = __ballot(k < n && st[k] == OVG_EMPTY);
    if (b && (threadIdx.x & 63) == __ffsll((long long)__ballot(1)) - 1)
        atomicMin(state + DT_LO, k - (int)(threadIdx.x & 63) + __ffsll((long long)b) - 1);
}

struct RemeshArgs {
    int n;  // slots
    const int *cand, *n_cand;
    int *gheap;       // heap entries past the LDS part
    int gheap_cap;
    int *fwd, *rev, *st;
    double *X, *Y, *Z;
    double *vs[3], *vels[3];
    const TangleParams *P;
    double min_dist, max_dist;
    int *state;
};

__device__ __forceinline__ void dt_pos(const RemeshArgs &a, int k, double &x, double &y,
                                       double &z)
{
    x = ld(a.X + k);
    y = ld(a.Y + k);
    z = ld(a.Z + k);
}

// update_tangent_normal (src/tangle.c:423-502) with the current links
__device__ void dt_stencil(const RemeshArgs &a, const TangleParams &P, int k, double t[3],
                           double nr[3])
{
    const int i1 = ld(a.fwd + k), im1 = ld(a.rev + k);
    const int i2 = ld(a.fwd + i1), im2 = ld(a.rev + im1);
    double p[5][3];
    const int s[5] = {k, i1, i2, im1, im2};
    for (int j = 0; j < 5; ++j) dt_pos(a, s[j], p[j][0], p[j][1], p[j][2]);
    tangent_normal_pts(P, p[0][0], p[0][1], p[0][2], p[1][0], p[1][1], p[1][2], p[2][0], p[2][1],
                       p[2][2], p[3][0], p[3][1], p[3][2], p[4][0], p[4][1], p[4][2], t, nr);
}

// The sequential remesh (src/tangle.c:1022-1056) over the candidate slots
// and the slots the pass's own edits reach ahead of its cursor, in slot
// order, as the host pass (ovg_remesh).  One wavefront; lanes share only
// the allocator's scan for an EMPTY slot.
__global__ __launch_bounds__(64) void k_dt_remesh(RemeshArgs a)
{
    // a binary min-heap whose first kHeapLds entries (its top levels) live
    // in LDS and the rest in global memory
    __shared__ int lheap[kHeapLds];
    const TangleParams &P = *a.P;
    const int lane = threadIdx.x;
    const int n = a.n, n_cand = *a.n_cand;
    int hsize = 0, hint = a.state[DT_HINT], lo = a.state[DT_LO];
    int cursor = -1, ci = 0, added = 0, removed = 0, err = 0;
    auto hget = [&](int i) { return i < kHeapLds ? lheap[i] : ld(a.gheap + (i - kHeapLds)); };
    auto hset = [&](int i, int v) {
        if (i < kHeapLds) lheap[i] = v;
        else a.gheap[i - kHeapLds] = v;
    };
    auto push = [&](int v) {
        if (hsize == kHeapLds + a.gheap_cap) {
            err |= kErrHeap;
            return;
        }
        int i = hsize++;
        while (i > 0) {
            const int p = (i - 1) >> 1;
            const int hp = hget(p);
            if (hp <= v) break;
            hset(i, hp);
            i = p;
        }
        hset(i, v);
    };
    auto pop = [&]() {
        const int v = hget(--hsize);
        int i = 0;
        for (;;) {
            int ch = 2 * i + 1;
            if (ch >= hsize) break;
            if (ch + 1 < hsize && hget(ch + 1) < hget(ch)) ++ch;
            const int hc = hget(ch);
            if (v <= hc) break;
            hset(i, hc);
            i = ch;
        }
        if (hsize > 0) hset(i, v);
    };
    auto touch = [&](int sl) {
        if (sl >= 0 && sl > cursor) push(sl);
    };
    // search_next_empty_node (src/tangle.c:201-234); slots below lo are
    // known not to be EMPTY, so the scan from slot 0 starts there
    auto alloc = [&]() -> int {
        if (hint > 0) {
            if (hint < n && ld(a.st + hint) == OVG_EMPTY) {
                const int r = hint;
                hint = (hint + 1 < n && ld(a.st + hint + 1) == OVG_EMPTY) ? hint + 1 : -1;
                return r;
            }
        }
        for (int base = lo; base < n; base += 64) {
            const int s = base + lane;
            const unsigned long long b = __ballot(s < n && ld(a.st + s) == OVG_EMPTY);
            if (b) {
                const int k = base + __ffsll((long long)b) - 1;
                hint = (k + 1 < n && ld(a.st + k + 1) == OVG_EMPTY) ? k + 1 : -1;
                lo = k + 1;
                return k;
            }
        }
        return -1;
    };
    for (;;) {
        const int av = ci < n_cand ? a.cand[ci] : INT_MAX;
        const int bv = hsize ? lheap[0] : INT_MAX;
        const int k = av < bv ? av : bv;
        if (k == INT_MAX || err) break;
        if (av == k) ++ci;
        while (hsize && lheap[0] == k) pop();
        if (k <= cursor) continue;
        cursor = k;
        if (k >= n || ld(a.st + k) == OVG_EMPTY) continue;
        const int next = ld(a.fwd + k), prev = ld(a.rev + k);
        double x, y, z;
        dt_pos(a, k, x, y, z);
        double lf = 0, lr = 0;
        if (next >= 0) {
            double bx, by, bz;
            dt_pos(a, next, bx, by, bz);
            lf = dt_dist(P, x, y, z, bx, by, bz);
        }
        if (prev >= 0) {
            double bx, by, bz;
            dt_pos(a, prev, bx, by, bz);
            lr = dt_dist(P, x, y, z, bx, by, bz);
        }
        if (next >= 0 && prev >= 0 &&
            ((lf < a.min_dist || lr < a.min_dist) && (lf + lr) < a.max_dist)) {
            // remove_point (:342-355)
            const int pv = ld(a.rev + k), nx = ld(a.fwd + k);
            if (pv >= 0) a.fwd[pv] = nx;
            if (nx >= 0) a.rev[nx] = pv;
            a.fwd[k] = -1;
            a.rev[k] = -1;
            a.st[k] = OVG_EMPTY;
            if (k < lo) lo = k;
            ++removed;
            touch(k);
            touch(pv);
            touch(nx);
        }
        if (next >= 0 && lf > a.max_dist) {
            // add_point (:261-335)
            const int nxt = ld(a.fwd + k);
            const int nw = alloc();
            if (nw < 0) {
                err |= kErrSlots;
                break;
            }
            a.st[nw] = OVG_FREE;
            const int p = k, nn = ld(a.fwd + p);
            double t0[3], n0[3], t1[3], n1[3];
            dt_stencil(a, P, p, t0, n0);
            dt_stencil(a, P, nn, t1, n1);
            double s0[3], s1[3];
            dt_pos(a, p, s0[0], s0[1], s0[2]);
            dt_pos(a, nn, s1[0], s1[1], s1[2]);
            pwrap(P, s0[0], s0[1], s0[2], s1[0], s1[1], s1[2]);
            const double l = dist3(s0[0], s0[1], s0[2], s1[0], s1[1], s1[2]);
            double nv[3], np[3];
            for (int q = 0; q < 3; ++q) nv[q] = n0[q] + n1[q];
            for (int q = 0; q < 3; ++q) nv[q] = nv[q] * 0.5;
            const double lnv = sqrt(dot3(nv[0], nv[1], nv[2], nv[0], nv[1], nv[2]));
            if (lnv > 1e-5) {
                const double R = 1 / lnv;
                const double dR = R * R - l * l / 4;
                const double delta = dR > 0 ? R - sqrt(dR) : R;
                for (int q = 0; q < 3; ++q) nv[q] = nv[q] * (1 / lnv);
                for (int q = 0; q < 3; ++q) nv[q] = nv[q] * -1;
                double aa[3], bb[3];
                for (int q = 0; q < 3; ++q) aa[q] = s0[q] + s1[q];
                for (int q = 0; q < 3; ++q) aa[q] = aa[q] * 0.5;
                for (int q = 0; q < 3; ++q) bb[q] = nv[q] * delta;
                for (int q = 0; q < 3; ++q) np[q] = aa[q] + bb[q];
            } else {
                for (int q = 0; q < 3; ++q) np[q] = s0[q] + s1[q];
                for (int q = 0; q < 3; ++q) np[q] = np[q] * 0.5;
            }
            a.X[nw] = np[0];
            a.Y[nw] = np[1];
            a.Z[nw] = np[2];
            for (int q = 0; q < 3; ++q) {
                a.vs[q][nw] = 0.0;
                a.vels[q][nw] = 0.0;
            }
            a.rev[nw] = p;
            a.fwd[nw] = nn;
            a.fwd[p] = nw;
            a.rev[nn] = nw;
            ++added;
            touch(nw);
            touch(k);
            touch(nxt);
        }
    }
    if (lane == 0) {
        a.state[DT_HINT] = hint;
        a.state[DT_ADDED] = added;
        a.state[DT_REMOVED] = removed;
        if (err) atomicOr(a.state + DT_ERR, err);
    }
}

// ------------------------------------------------------------ reconnection
struct PairRec {
    int k, l, ek, el;   // slots, endpoint indices
    double dx[3];       // x_k - seg_pwrap(x_k, x_l).r2
    int info;           // bits 0-1: do_reconnection's case with the links before the
                        // pass (0 none); bit 2: the tangent tests reject the pair
                        // with the endpoints' tangents before the pass
    int chg;            // first of its two changes once rewired, else -1
};

struct EndRec {
    int slot, fwd, rev, wf, wr;  // slot, its links before the pass, their endpoint indices
    int near_n;                  // tree mode: rewirable pairs near it (-1: more than kNear)
    int applied;                 // direct mode: changes added into bs
    int pad;
    int near[kNear];
    double t[3], nr[3];          // tangent, normal before the pass
    double f_lia;                // lia_velocity's factor before the pass
    double bs[3];                // Biot-Savart before the pass (direct mode: plus the
                                 // field changes of the first `applied` rewirings)
    double vel[3];               // velocity composed from the values before the pass
};

// (k, l) -> sortable key; slots past the pass's pairs sort last
__global__ void k_dt_pair_keys(int cap, int bits, const int2 *__restrict__ pairs,
                               const int *__restrict__ found, uint64_t *__restrict__ keys,
                               int *__restrict__ vals, int *__restrict__ state)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int cnt = *found;
    const int np = cnt <= cap ? cnt : 0;
    if (i == 0) {
        state[DT_FOUND] = cnt;
        state[DT_PAIRS] = np;
        if (cnt > cap) atomicOr(state + DT_ERR, kErrPairs);
    }
    if (i >= cap) return;
    if (i < np) {
        const int2 p = pairs[i];
        keys[i] = ((uint64_t)(unsigned)p.x << bits) | (uint64_t)(unsigned)p.y;
    } else {
        keys[i] = ~0ull;
    }
    vals[i] = i;
}

__global__ void k_dt_pair_unpack(int cap, int bits, const uint64_t *__restrict__ keys,
                                 const int *__restrict__ state, PairRec *__restrict__ pr,
                                 int *__restrict__ mark)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cap || j >= state[DT_PAIRS]) return;
    const uint64_t key = keys[j];
    const int k = (int)(key >> bits), l = (int)(key & ((1ull << bits) - 1));
    pr[j].k = k;
    pr[j].l = l;
    pr[j].chg = -1;
    mark[k] = 1;
    mark[l] = 1;
}

// endpoints in slot order; where[slot] = index.  mark returns to zero.
__global__ void k_dt_ends(int n, int bound, int *__restrict__ mark, const int *__restrict__ pos,
                          const int *__restrict__ total, EndRec *__restrict__ er,
                          int *__restrict__ eslot, int *__restrict__ where,
                          int *__restrict__ state)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = *total;
    if (s == 0) {
        state[DT_ENDS] = m <= bound ? m : 0;
        if (m > bound) atomicOr(state + DT_ERR, kErrEnds);
    }
    if (s >= n || !mark[s]) return;
    mark[s] = 0;
    if (m > bound) return;
    const int e = pos[s];
    er[e].slot = s;
    eslot[e] = s;
    where[s] = e;
}

__global__ void k_dt_end_prep(int bound, const int *__restrict__ state, EndRec *__restrict__ er,
                              const int *__restrict__ where, const double *__restrict__ X,
                              const double *__restrict__ Y, const double *__restrict__ Z,
                              const int *__restrict__ fwd, const int *__restrict__ rev,
                              const int *__restrict__ st, const TangleParams *__restrict__ Pp)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= bound || e >= state[DT_ENDS]) return;
    const TangleParams &P = *Pp;
    EndRec &E = er[e];
    const int s = E.slot, f = fwd[s], r = rev[s];
    E.fwd = f;
    E.rev = r;
    E.wf = where[f];
    E.wr = where[r];
    E.applied = 0;
    E.near_n = 0;
    double t[3], nr[3];
    tangent_normal(P, X, Y, Z, fwd, rev, st, nullptr, s, t, nr);
    for (int q = 0; q < 3; ++q) {
        E.t[q] = t[q];
        E.nr[q] = nr[q];
    }
    // lia_velocity's factor (src/tangle.c:568-592) with these links, in the
    // host composition's order
    const double px = X[s], py = Y[s], pz = Z[s];
    const double l_next = dt_dist(P, px, py, pz, X[f], Y[f], Z[f]);
    double wx = px, wy = py, wz = pz;
    const double qx = X[r], qy = Y[r], qz = Z[r];
    pwrap(P, qx, qy, qz, wx, wy, wz);
    const double l_prev = dist3(qx, qy, qz, wx, wy, wz);
    E.f_lia = P.kappa * log(2 * sqrt(l_next * l_prev) / P.sqrtE / P.a0) / 4 / M_PI;
}

// the composition of src/tangle.c:600-726 for a FREE node, as the host
// pass composes it (host_phys.h compose): vs = LIA + v_s + bs, then mutual
// friction
__device__ __forceinline__ void dt_compose(const TangleParams &P, double f, const double t[3],
                                           const double nr[3], const double bs[3], double v[3])
{
    const double c0 = t[1] * nr[2] - t[2] * nr[1];
    const double c1 = t[2] * nr[0] - t[0] * nr[2];
    const double c2 = t[0] * nr[1] - t[1] * nr[0];
    double vs[3] = {c0 * f, c1 * f, c2 * f};
    for (int q = 0; q < 3; ++q) vs[q] = vs[q] + P.evs[q];
    for (int q = 0; q < 3; ++q) vs[q] = vs[q] + bs[q];
    for (int q = 0; q < 3; ++q) v[q] = vs[q];
    if (P.use_mf) {
        double dv[3];
        for (int q = 0; q < 3; ++q) dv[q] = P.evn[q] - vs[q];
        const double m0 = t[1] * dv[2] - t[2] * dv[1];
        const double m1 = t[2] * dv[0] - t[0] * dv[2];
        const double m2 = t[0] * dv[1] - t[1] * dv[0];
        v[0] = v[0] + m0 * P.alpha;
        v[1] = v[1] + m1 * P.alpha;
        v[2] = v[2] + m2 * P.alpha;
        const double n0 = t[1] * m2 - t[2] * m1;
        const double n1 = t[2] * m0 - t[0] * m2;
        const double n2 = t[0] * m1 - t[1] * m0;
        v[0] = v[0] + n0 * -P.alpha_p;
        v[1] = v[1] + n1 * -P.alpha_p;
        v[2] = v[2] + n2 * -P.alpha_p;
    }
}

__global__ void k_dt_end_vel(int bound, const int *__restrict__ state, EndRec *__restrict__ er,
                             const double *__restrict__ bpx, const double *__restrict__ bpy,
                             const double *__restrict__ bpz, const double *__restrict__ bix,
                             const double *__restrict__ biy, const double *__restrict__ biz,
                             const TangleParams *__restrict__ Pp)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= bound || e >= state[DT_ENDS]) return;
    EndRec &E = er[e];
    const int s = E.slot;
    const double bs[3] = {bpx[s] + bix[s], bpy[s] + biy[s], bpz[s] + biz[s]};
    double t[3], nr[3], v[3];
    for (int q = 0; q < 3; ++q) {
        E.bs[q] = bs[q];
        t[q] = E.t[q];
        nr[q] = E.nr[q];
    }
    dt_compose(*Pp, E.f_lia, t, nr, bs, v);
    for (int q = 0; q < 3; ++q) E.vel[q] = v[q];
}

__device__ __forceinline__ double len3v(const double *u)
{
    return sqrt(dot3(u[0], u[1], u[2], u[0], u[1], u[2]));
}

// the tangent tests of src/vortex_dynamics.c:236-246
__device__ __forceinline__ bool dt_angle_reject(const double *tk, const double *tl, double cutoff)
{
    if (len3v(tk) < 0.5 || len3v(tl) < 0.5) return true;
    const double ca = dot3(tk[0], tk[1], tk[2], tl[0], tl[1], tl[2]) / len3v(tk) / len3v(tl);
    return ca > cutoff;
}

__global__ void k_dt_pair_prep(int cap, const int *__restrict__ state, PairRec *__restrict__ pr,
                               const EndRec *__restrict__ er, const int *__restrict__ where,
                               const double *__restrict__ X, const double *__restrict__ Y,
                               const double *__restrict__ Z, const TangleParams *__restrict__ Pp,
                               double cutoff)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cap || j >= state[DT_PAIRS] || state[DT_ENDS] == 0) return;
    const TangleParams &P = *Pp;
    PairRec &R = pr[j];
    const int k = R.k, l = R.l, ek = where[k], el = where[l];
    R.ek = ek;
    R.el = el;
    const double xk = X[k], yk = Y[k], zk = Z[k];
    double wx = X[l], wy = Y[l], wz = Z[l];
    pwrap(P, xk, yk, zk, wx, wy, wz);
    R.dx[0] = xk - wx;
    R.dx[1] = yk - wy;
    R.dx[2] = zk - wz;
    const EndRec &K = er[ek], &L = er[el];
    const bool rej = dt_angle_reject(K.t, L.t, cutoff);
    // do_reconnection's test (:128-152) with the links before the pass,
    // which an unflagged pair still has when the greedy pass reaches it
    const int kf = K.fwd, kr = K.rev, lf = L.fwd, lr = L.rev;
    auto plen = [&](int a, int b) {
        return dt_dist(P, X[a], Y[a], Z[a], X[b], Y[b], Z[b]);
    };
    const double l_kl = plen(k, l), l_kflr = plen(kf, lr), l_kkf = plen(k, kf),
                 l_llr = plen(l, lr);
    const double l_krlf = plen(kr, lf), l_kkr = plen(k, kr), l_llf = plen(l, lf);
    const double cf1 = l_kl + l_kflr - l_kkf - l_llr;
    const double cf2 = l_kl + l_krlf - l_kkr - l_llf;
    const int kind = (cf1 > 0 && cf2 > 0) ? 0 : (cf1 < cf2 ? 1 : 2);
    R.info = kind | (rej ? 4 : 0);
}

// |a - b|^2 with periodic axes folded to the nearest image, the host pass's
// `near` test
__device__ __forceinline__ double dt_near2(const TangleParams &P, double ax, double ay, double az,
                                           double bx, double by, double bz)
{
    double d[3] = {ax - bx, ay - by, az - bz};
    double d2 = 0;
    for (int q = 0; q < 3; ++q) {
        if (P.periodic[q]) d[q] -= P.L[q] * rint(d[q] / P.L[q]);
        d2 += d[q] * d[q];
    }
    return d2;
}

// Tree mode: for every endpoint, the pairs whose rewiring could change a
// segment within the correction radius of it (a rewired segment starts at
// k, l or their reverse neighbours), ascending; more than kNear: -1.
__global__ void k_dt_near(int bound, const int *__restrict__ state, EndRec *__restrict__ er,
                          const PairRec *__restrict__ pr, const double *__restrict__ X,
                          const double *__restrict__ Y, const double *__restrict__ Z,
                          const TangleParams *__restrict__ Pp, double rc2)
{
    __shared__ int cnt;
    __shared__ int lst[kNear];
    const int e = blockIdx.x;
    if (e >= bound || e >= state[DT_ENDS]) return;
    const TangleParams &P = *Pp;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int s = er[e].slot;
    const double px = X[s], py = Y[s], pz = Z[s];
    const int np = state[DT_PAIRS];
    const double lim = rc2 * (1 + 1e-9);
    for (int j = threadIdx.x; j < np; j += blockDim.x) {
        const PairRec &R = pr[j];
        const EndRec &K = er[R.ek], &L = er[R.el];
        const int segs[4] = {R.k, R.l, K.rev, L.rev};
        bool hit = false;
        for (int q = 0; q < 4 && !hit; ++q) {
            const int g = segs[q];
            if (g >= 0) hit = dt_near2(P, X[g], Y[g], Z[g], px, py, pz) < lim;
        }
        if (hit) {
            const int at = atomicAdd(&cnt, 1);
            if (at < kNear) lst[at] = j;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int c = cnt;
        if (c > kNear) {
            er[e].near_n = -1;
        } else {
            for (int i = 1; i < c; ++i)  // ascending pair order
                for (int q = i; q > 0 && lst[q - 1] > lst[q]; --q) {
                    const int tmp = lst[q];
                    lst[q] = lst[q - 1];
                    lst[q - 1] = tmp;
                }
            er[e].near_n = c;
            for (int i = 0; i < c; ++i) er[e].near[i] = lst[i];
        }
    }
}

// Target groups of the endpoints for the list traversal, as the host path
// (launch_bh_list_targets) forms them: endpoints in tree order, a new group
// after points_per_wave targets or where two keys share fewer than the
// group level's digits.  More endpoints than one LDS sort holds: one group
// each.
__global__ __launch_bounds__(1024) void k_dt_groups(
    int bound, const int *__restrict__ state, const int *__restrict__ eslot,
    const int *__restrict__ rank, const uint64_t *__restrict__ keys, int ppw, int L0,
    int2 *__restrict__ grp, int *__restrict__ tgt, int *__restrict__ n_groups)
{
    __shared__ uint64_t sk[kGroupSort];
    const int m = min(state[DT_ENDS], bound);
    const int t = threadIdx.x;
    if (m > kGroupSort) {
        for (int i = t; i < m; i += blockDim.x) {
            tgt[i] = eslot[i];
            grp[i] = make_int2(i, 1);
        }
        if (t == 0) *n_groups = m;
        return;
    }
    int sz = 1;
    while (sz < m) sz <<= 1;
    for (int i = t; i < sz; i += blockDim.x) {
        if (i < m) {
            const int s = eslot[i];
            sk[i] = ((uint64_t)(unsigned)rank[s] << 32) | (unsigned)s;
        } else {
            sk[i] = ~0ull;
        }
    }
    __syncthreads();
    for (int k = 2; k <= sz; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < sz; i += blockDim.x) {
                const int p = i ^ j;
                if (p > i) {
                    const uint64_t a = sk[i], b = sk[p];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        sk[i] = b;
                        sk[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    if (t == 0) {
        int ng = 0, cur = 0;
        uint64_t prev = 0;
        for (int i = 0; i < m; ++i) {
            const int r = (int)(sk[i] >> 32);
            const uint64_t ky = keys[r];
            bool cut = i == 0 || cur == ppw;
            if (!cut) {
                // common 3-bit levels of one tangle's 48-bit Morton keys
                const int lv = ky == prev ? 16 : (__clzll((long long)(prev ^ ky)) - 15) / 3;
                cut = lv < L0;
            }
            if (cut) {
                grp[ng++] = make_int2(i, 1);
                cur = 1;
            } else {
                grp[ng - 1].y++;
                ++cur;
            }
            tgt[i] = (int)(sk[i] & 0xffffffffu);
            prev = ky;
        }
        *n_groups = ng;
    }
}

struct GreedyArgs {
    PairRec *pr;
    EndRec *er;
    const int *where;
    int *fwd, *rev;
    const double *X, *Y, *Z;
    int4 *chg;       // (segment start, old forward, new forward, pair)
    const TangleParams *P;
    double cutoff, rc2;
    int exact, max_chg;
    int *state;
};

constexpr int kGreedyWords = kMaxEnds / 32;

// The field of segment (m -> f) at slot p, all images (host_phys.h
// segment_all_images with reflect = false: pure shift images), in image
// order; primary box skipped when p is an end of the segment.
__device__ void dt_segment_all(const GreedyArgs &a, const TangleParams &P, int m, int f, int p,
                               double acc[3])
{
    const double xm = a.X[m], ym = a.Y[m], zm = a.Z[m];
    double bx = a.X[f], by = a.Y[f], bz = a.Z[f];
    pwrap(P, xm, ym, zm, bx, by, bz);
    const double rx = a.X[p], ry = a.Y[p], rz = a.Z[p];
    for (int im = -1; im < P.n_images; ++im) {
        if (im < 0 && (m == p || f == p)) continue;
        double qx = rx, qy = ry, qz = rz;
        if (im >= 0) image_point(P, im, rx, ry, rz, qx, qy, qz);
        double vx, vy, vz;
        if (seg_field64(xm - qx, ym - qy, zm - qz, bx - qx, by - qy, bz - qz, P.f_bs, vx, vy,
                        vz)) {
            acc[0] += vx;
            acc[1] += vy;
            acc[2] += vz;
        }
    }
}

// The sequential greedy pass of src/vortex_dynamics.c:224-284 in (k, l)
// order, as host_recon.h greedy_pass: one wavefront.
__global__ __launch_bounds__(64) void k_dt_greedy(GreedyArgs a)
{
    __shared__ unsigned fbits[kGreedyWords], sbits[kGreedyWords];
    __shared__ double sd[64][3];
    __shared__ int sfound[64];
    const TangleParams &P = *a.P;
    const int lane = threadIdx.x;
    const int np = a.state[DT_PAIRS], m = a.state[DT_ENDS];
    if (np == 0 || m == 0) {
        if (lane == 0) {
            a.state[DT_REC] = 0;
            a.state[DT_CHANGES] = 0;
        }
        return;
    }
    for (int i = lane; i < kGreedyWords; i += 64) {
        fbits[i] = 0;
        sbits[i] = 0;
    }
    __syncthreads();
    auto bit = [&](const unsigned *b, int e) { return e >= 0 && ((b[e >> 5] >> (e & 31)) & 1); };
    auto set = [&](unsigned *b, int e) {
        if (e >= 0) b[e >> 5] |= 1u << (e & 31);
    };
    int nch = 0, nrec = 0;
    // tangent and normal of endpoint e: before the pass unless a rewiring
    // changed a link of its stencil, else from the current links
    auto tangent = [&](int e, double t[3], double nr[3]) {
        const EndRec &E = a.er[e];
        if (!bit(sbits, e)) {
            for (int q = 0; q < 3; ++q) {
                t[q] = E.t[q];
                nr[q] = E.nr[q];
            }
            return;
        }
        const int s = E.slot;
        const int i1 = ld(a.fwd + s), im1 = ld(a.rev + s);
        const int i2 = ld(a.fwd + i1), im2 = ld(a.rev + im1);
        tangent_normal_pts(P, a.X[s], a.Y[s], a.Z[s], a.X[i1], a.Y[i1], a.Z[i1], a.X[i2],
                           a.Y[i2], a.Z[i2], a.X[im1], a.Y[im1], a.Z[im1], a.X[im2], a.Y[im2],
                           a.Z[im2], t, nr);
    };
    // bs += (field of the new segment - field of the old one), change by
    // change in order, for the cnt changes listed in sfound
    auto apply = [&](int cnt, int p, double bs[3]) {
        for (int b0 = 0; b0 < cnt; b0 += 64) {
            const int nb = min(64, cnt - b0);
            __syncthreads();
            if (lane < nb) {
                const int4 ch = a.chg[sfound[b0 + lane]];
                double vo[3] = {0, 0, 0}, vn[3] = {0, 0, 0};
                dt_segment_all(a, P, ch.x, ch.y, p, vo);
                dt_segment_all(a, P, ch.x, ch.z, p, vn);
                for (int q = 0; q < 3; ++q) sd[lane][q] = vn[q] - vo[q];
            }
            __syncthreads();
            for (int i = 0; i < nb; ++i)
                for (int q = 0; q < 3; ++q) bs[q] += sd[i][q];
        }
    };
    auto velocity = [&](int e, const double t[3], const double nr[3], double v[3]) {
        EndRec &E = a.er[e];
        const int p = E.slot;
        double bs[3] = {ld(&E.bs[0]), ld(&E.bs[1]), ld(&E.bs[2])};
        bool found = false;
        if (a.exact) {
            // every rewiring so far, added once into the endpoint's running sum
            found = nch > 0;
            int app = ld(&E.applied);
            while (app < nch) {
                const int nb = min(64, nch - app);
                __syncthreads();
                if (lane < nb) sfound[lane] = app + lane;
                __syncthreads();
                apply(nb, p, bs);
                app += nb;
            }
            if (ld(&E.applied) != app) {
                E.bs[0] = bs[0];
                E.bs[1] = bs[1];
                E.bs[2] = bs[2];
                E.applied = app;
            }
        } else {
            // the rewirings within the correction radius, in rewiring order
            const double px = a.X[p], py = a.Y[p], pz = a.Z[p];
            int cnt = 0;
            __syncthreads();
            if (E.near_n >= 0) {
                for (int i = 0; i < E.near_n; ++i) {
                    const int c0 = ld(&a.pr[E.near[i]].chg);
                    if (c0 < 0) continue;
                    for (int c = c0; c < c0 + 2; ++c) {
                        const int g = ld(&a.chg[c].x);
                        if (dt_near2(P, a.X[g], a.Y[g], a.Z[g], px, py, pz) < a.rc2) {
                            if (lane == 0) sfound[cnt] = c;
                            ++cnt;
                        }
                    }
                }
                __syncthreads();
                apply(cnt, p, bs);
            } else {
                for (int c0 = 0; c0 < nch; c0 += 64) {
                    const int c = c0 + lane;
                    bool hit = false;
                    if (c < nch) {
                        const int g = ld(&a.chg[c].x);
                        hit = dt_near2(P, a.X[g], a.Y[g], a.Z[g], px, py, pz) < a.rc2;
                    }
                    const unsigned long long b = __ballot(hit);
                    const int nb = __popcll(b);
                    if (nb == 0) continue;
                    __syncthreads();
                    if (hit) sfound[__popcll(b & ((1ull << lane) - 1))] = c;
                    __syncthreads();
                    apply(nb, p, bs);
                    cnt += nb;
                }
            }
            found = cnt > 0;
        }
        if (!found && !bit(sbits, e)) {
            for (int q = 0; q < 3; ++q) v[q] = E.vel[q];
            return;
        }
        dt_compose(P, E.f_lia, t, nr, bs, v);
    };

    int cur_k = -1;
    bool skip = false, have_t = false, have_v = false;
    double tk[3], nk[3], vk[3];
    for (int j = 0; j < np; ++j) {
        PairRec &R = a.pr[j];
        const int k = R.k, l = R.l, ek = R.ek, el = R.el;
        if (k != cur_k) {
            cur_k = k;
            skip = bit(fbits, ek);
            have_t = false;
            have_v = false;
        }
        if (skip || bit(fbits, el)) continue;
        const int info = R.info;
        double tl[3], nl[3];
        bool have_tl = false;
        if (!bit(sbits, ek) && !bit(sbits, el)) {
            if (info & 4) continue;
        } else {
            if (!have_t) {
                tangent(ek, tk, nk);
                have_t = true;
            }
            tangent(el, tl, nl);
            have_tl = true;
            if (dt_angle_reject(tk, tl, a.cutoff)) continue;
        }
        if (!have_v) {
            if (!have_t) {
                tangent(ek, tk, nk);
                have_t = true;
            }
            velocity(ek, tk, nk, vk);
            have_v = true;
        }
        if (!have_tl) tangent(el, tl, nl);
        double vl[3];
        velocity(el, tl, nl, vl);
        double dv[3];
        for (int q = 0; q < 3; ++q) dv[q] = vk[q] - vl[q];
        if (dot3(R.dx[0], R.dx[1], R.dx[2], dv[0], dv[1], dv[2]) > 0) continue;
        const int kind = info & 3;
        if (!kind) continue;
        // do_reconnection's rewiring (:153-172) and the changed segments
        const EndRec &K = a.er[ek], &L = a.er[el];
        const int kf = K.fwd, kr = K.rev, lf = L.fwd, lr = L.rev;
        if (nch + 2 > a.max_chg) break;  // cannot happen: two per pair
        int4 c0, c1;
        if (kind == 1) {
            a.fwd[k] = l;
            a.rev[l] = k;
            a.rev[kf] = lr;
            a.fwd[lr] = kf;
            c0 = make_int4(k, kf, l, j);
            c1 = make_int4(lr, l, kf, j);
        } else {
            a.fwd[l] = k;
            a.rev[k] = l;
            a.rev[lf] = kr;
            a.fwd[kr] = lf;
            c0 = make_int4(l, lf, k, j);
            c1 = make_int4(kr, k, lf, j);
        }
        a.chg[nch] = c0;
        a.chg[nch + 1] = c1;
        R.chg = nch;
        nch += 2;
        ++nrec;
        // flags: the pair, its old neighbours (and the new ones, which are
        // among them)
        set(fbits, ek);
        set(fbits, el);
        set(fbits, K.wf);
        set(fbits, K.wr);
        set(fbits, L.wf);
        set(fbits, L.wr);
        // stale stencils: the six slots and their current neighbours
        const int six[6] = {k, l, kf, kr, lf, lr};
        for (int q = 0; q < 6; ++q) {
            const int sl = six[q];
            set(sbits, a.where[sl]);
            const int f = ld(a.fwd + sl), r = ld(a.rev + sl);
            if (f >= 0) set(sbits, a.where[f]);
            if (r >= 0) set(sbits, a.where[r]);
        }
        skip = true;
    }
    if (lane == 0) {
        a.state[DT_REC] = nrec;
        a.state[DT_CHANGES] = nch;
    }
}

__global__ void k_dt_clear_where(int bound, const int *__restrict__ state,
                                 const EndRec *__restrict__ er, int *__restrict__ where)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= bound || e >= state[DT_ENDS]) return;
    where[er[e].slot] = -1;
}

// ======================================================================
// host side

static inline int nb256(int n) { return (n + 255) / 256; }

static int dt_state_ready(Ctx &c)
{
    if (c.dt_state.p) return OVG_OK;
    OVG_TRY(c.dt_state.ensure(sizeof(int) * DT_COUNT));
    hipLaunchKernelGGL(k_dt_fill, dim3(1), dim3(64), 0, c.stream, DT_COUNT, c.dt_state.as<int>(),
                       0);
    return OVG_OK;
}

bool devtopo_applies(const Ctx &c)
{
    if (c.params.topology != OVG_TOPO_DEVICE || c.n_tangles > 1 || c.has_pinned) return false;
    for (int a = 0; a < 6; ++a)
        if (c.params.wall[a] != OVG_WALL_PERIODIC) return false;
    for (int j = 0; j < c.params.n_images; ++j)
        if (c.params.image_reflect[j]) return false;
    // tree-mode reconnection velocities come from the list traversal
    if (c.params.rec_velocity == OVG_RECVEL_TREE && c.params.method == OVG_METHOD_BH &&
        c.params.traverse != OVG_TRAVERSE_LIST)
        return false;
    return true;
}

// after a device pass edited links, status or positions
static void device_edited(Ctx &c)
{
    c.dt_unread = true;
    c.mirrors_stale = true;
    c.rec_ok = false;  // recorded interaction lists follow the old links
    c.loops_full_scan = true;
    c.topo_version++;
}

int sync_live(Ctx &c)
{
    if (!c.dt_unread) return OVG_OK;
    hipSetDevice(c.device);
    if (c.live_pending) OVG_TRY(rebuild_live(c, c.dt_state.as<int>() + DT_LIVE));
    int st[DT_COUNT];
    OVG_HIP(hipMemcpyAsync(st, c.dt_state.p, sizeof(st), hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    c.dt_unread = false;
    if (c.live_pending) {
        c.live_pending = false;
        c.n_live = st[DT_LIVE];
        c.n_src = st[DT_LIVE];
    }
    c.dt_seen = std::max(c.dt_seen, st[DT_FOUND]);
    c.dt_ends_seen = std::max(c.dt_ends_seen, st[DT_ENDS]);
    if (st[DT_ERR]) {
        hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, c.dt_state.as<int>(), DT_ERR,
                           0);
        const int e = st[DT_ERR];
        if (e & kErrPairs)
            return set_error(OVG_ELIMIT,
                             "device reconnection pass: %d candidate pairs exceed its capacity "
                             "of %d; the pass did not run (the capacity grows for later passes, "
                             "up to %d)",
                             st[DT_FOUND], c.dt_cap, kMaxPairs);
        if (e & kErrEnds)
            return set_error(OVG_ELIMIT,
                             "device reconnection pass: more than %d endpoints; the pass did "
                             "not run",
                             kMaxEnds);
        if (e & kErrSlots)
            return set_error(OVG_ELIMIT,
                             "device remesh: no EMPTY slot left among %d for an insertion; the "
                             "pass stopped there",
                             c.n);
        return set_error(OVG_ELIMIT,
                         "device remesh: its heap of touched slots overflowed; the pass "
                         "stopped there");
    }
    return OVG_OK;
}

int sync_mirrors(Ctx &c)
{
    OVG_TRY(sync_live(c));
    if (!c.mirrors_stale) return OVG_OK;
    hipSetDevice(c.device);
    const int n = c.n;
    OVG_TRY(c.partial.ensure(sizeof(int) * 3 * (size_t)n + 64));
    int *h = nullptr;
    std::vector<int> buf(3 * (size_t)n + 1);
    h = buf.data();
    OVG_HIP(hipMemcpyAsync(h, c.fwd.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost,
                           c.stream));
    OVG_HIP(hipMemcpyAsync(h + n, c.rev.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost,
                           c.stream));
    OVG_HIP(hipMemcpyAsync(h + 2 * (size_t)n, c.status.p, sizeof(int) * (size_t)n,
                           hipMemcpyDeviceToHost, c.stream));
    int hint = c.next_empty;
    if (c.hint_on_device)
        OVG_HIP(hipMemcpyAsync(&hint, c.dt_state.as<int>() + DT_HINT, sizeof(int),
                               hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    c.h_conn.resize(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        c.h_conn[2 * (size_t)k] = h[k];
        c.h_conn[2 * (size_t)k + 1] = h[n + k];
        // device mode holds no pinned nodes: pin_wall is NOT_A_FACE
        c.h_status[2 * (size_t)k] = h[2 * (size_t)n + k];
        c.h_status[2 * (size_t)k + 1] = -1;
    }
    c.next_empty = hint;
    c.hint_on_device = false;
    c.mirrors_stale = false;
    return OVG_OK;
}

int devtopo_counts(Ctx &c, int out[4])
{
    OVG_TRY(dt_state_ready(c));
    OVG_TRY(sync_live(c));
    int st[DT_COUNT];
    OVG_HIP(hipMemcpyAsync(st, c.dt_state.p, sizeof(st), hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    out[0] = st[DT_REC];
    out[1] = st[DT_LOOPS];
    out[2] = st[DT_ADDED];
    out[3] = st[DT_REMOVED];
    return OVG_OK;
}

int devtopo_small_loops(Ctx &c, int cutoff)
{
    OVG_TRY(dt_state_ready(c));
    const int n = c.n;
    int *state = c.dt_state.as<int>();
    hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_LOOPS, 0);
    if (n == 0) return OVG_OK;
    OVG_TRY(c.rc_tmp.ensure(sizeof(int) * (size_t)(n + 1)));
    int *flag = c.rc_tmp.as<int>();
    hipLaunchKernelGGL(k_dt_loop_flag, dim3(nb256(n)), dim3(256), 0, c.stream, n,
                       c.fwd.as<int>(), c.status.as<int>(), cutoff, flag);
    hipLaunchKernelGGL(k_dt_loop_erase, dim3(nb256(n)), dim3(256), 0, c.stream, n, flag,
                       c.fwd.as<int>(), c.rev.as<int>(), c.status.as<int>(), state);
    device_edited(c);
    c.live_pending = true;
    return launch_status();
}

int devtopo_remesh(Ctx &c, double min_dist, double max_dist)
{
    OVG_TRY(dt_state_ready(c));
    int *state = c.dt_state.as<int>();
    // room for the insertions: the reference doubles the slots when none is
    // EMPTY (get_next_empty_node, src/tangle.c:241-253); doubling earlier
    // leaves every allocation the same slot, since the scan finds the old
    // slots first either way.  The live count here may be an over-estimate.
    const int free_slots = c.n - c.n_live;
    if (free_slots < std::max(4096, c.n_live / 4)) {
        OVG_TRY(sync_mirrors(c));
        int grown = 2 * c.n;
        while (grown - c.n_live < std::max(4096, c.n_live / 4)) grown *= 2;
        OVG_TRY(grow_capacity(c, grown));
    }
    const int n = c.n;
    hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_ADDED, 0);
    hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_REMOVED, 0);
    if (n == 0) return OVG_OK;
    if (!c.hint_on_device) {
        hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_HINT,
                           c.next_empty);
        c.hint_on_device = true;
    }
    OVG_TRY(positions_change(c));
    const int *cand = nullptr, *n_cand = nullptr;
    OVG_TRY(remesh_candidates_queue(c, min_dist, max_dist, &cand, &n_cand));
    hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_LO, n);
    hipLaunchKernelGGL(k_dt_first_empty, dim3(nb256(n)), dim3(256), 0, c.stream, n,
                       c.status.as<int>(), state);
    // each visit touches at most 6 slots ahead of the cursor, and every
    // slot is visited at most once
    const size_t gcap = 6 * (size_t)n + 64;
    OVG_TRY(c.dt_chg.ensure(std::max(c.dt_chg.bytes, sizeof(int) * gcap)));
    RemeshArgs a{};
    a.n = n;
    a.cand = cand;
    a.n_cand = n_cand;
    a.gheap = c.dt_chg.as<int>();
    a.gheap_cap = (int)std::min<size_t>(gcap, INT32_MAX - kHeapLds);
    a.fwd = c.fwd.as<int>();
    a.rev = c.rev.as<int>();
    a.st = c.status.as<int>();
    a.X = c.x[0].as<double>();
    a.Y = c.x[1].as<double>();
    a.Z = c.x[2].as<double>();
    for (int q = 0; q < 3; ++q) {
        a.vs[q] = c.vs[q].as<double>();
        a.vels[q] = c.vels[q].as<double>();
    }
    a.P = dparams(c);
    a.min_dist = min_dist;
    a.max_dist = max_dist;
    a.state = state;
    hipLaunchKernelGGL(k_dt_remesh, dim3(1), dim3(64), 0, c.stream, a);
    // enforce_periodic_boundaries after insertions (:1057-1058); walls are
    // periodic here, so enforce_wall_boundaries changes nothing
    launch_wrap_slots(c, state + DT_ADDED);
    device_edited(c);
    c.live_pending = true;
    return launch_status();
}

int devtopo_reconnect(Ctx &c, double t)
{
    OVG_TRY(sync_live(c));
    OVG_TRY(dt_state_ready(c));
    int *state = c.dt_state.as<int>();
    hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_REC, 0);
    const int n = c.n;
    if (c.n_live == 0 || n == 0) return OVG_OK;
    // capacity: four times the most pairs seen, at least 4096
    if (c.dt_cap == 0) c.dt_cap = 8192;
    while (c.dt_cap < kMaxPairs && c.dt_cap < 4 * c.dt_seen) c.dt_cap *= 2;
    if (c.dt_cap > kMaxPairs) c.dt_cap = kMaxPairs;
    const int cap = c.dt_cap;
    const int ebound = std::min(2 * cap, kMaxEnds);
    c.rc_cap = cap;
    push_params(c, t);
    OVG_TRY(candidates_queue(c));
    // (k, l) order: one radix sort of packed keys
    int bits = 1;
    while ((1LL << bits) < (long long)n) ++bits;
    int passes = (2 * bits + 7) / 8;
    if (passes & 1) ++passes;
    TreeBufs &T = c.tree;
    for (int b = 0; b < 2; ++b) {
        OVG_TRY(T.keys[b].ensure(sizeof(uint64_t) * (size_t)cap));
        OVG_TRY(T.vals[b].ensure(sizeof(int) * (size_t)cap));
    }
    ++c.tree_id;
    OVG_TRY(c.dt_pairs.ensure(sizeof(PairRec) * (size_t)cap));
    OVG_TRY(c.dt_ends.ensure(sizeof(EndRec) * (size_t)ebound + sizeof(int) * (size_t)ebound + 256));
    OVG_TRY(c.dt_chg.ensure(sizeof(int4) * 2 * (size_t)cap + 64));
    OVG_TRY(c.dt_mark.ensure(sizeof(int) * (size_t)(n + 1)));
    OVG_TRY(c.dt_where.ensure(sizeof(int) * (size_t)(n + 1)));
    OVG_TRY(c.rc_tmp.ensure(sizeof(int) * (size_t)(n + 1)));
    PairRec *pr = c.dt_pairs.as<PairRec>();
    EndRec *er = c.dt_ends.as<EndRec>();
    int *eslot = reinterpret_cast<int *>(er + ebound);
    int *mark = c.dt_mark.as<int>(), *where = c.dt_where.as<int>(), *pos = c.rc_tmp.as<int>();
    // mark is zero and where -1 between passes; both are filled here
    // because a growth may have replaced them
    hipLaunchKernelGGL(k_dt_fill, dim3(nb256(n)), dim3(256), 0, c.stream, n, mark, 0);
    hipLaunchKernelGGL(k_dt_fill, dim3(nb256(n)), dim3(256), 0, c.stream, n, where, -1);
    hipLaunchKernelGGL(k_dt_pair_keys, dim3(nb256(cap)), dim3(256), 0, c.stream, cap, bits,
                       c.rc_pairs.as<int2>(), c.rc_count.as<int>(), T.keys[0].as<uint64_t>(),
                       T.vals[0].as<int>(), state);
    OVG_TRY(radix_sort(c, cap, 0, 8 * (passes - 1)));
    hipLaunchKernelGGL(k_dt_pair_unpack, dim3(nb256(cap)), dim3(256), 0, c.stream, cap, bits,
                       T.keys[0].as<uint64_t>(), state, pr, mark);
    OVG_TRY(launch_exclusive_scan(c, mark, pos, n, pos + n, c.tree.scan_tmp));
    hipLaunchKernelGGL(k_dt_ends, dim3(nb256(n)), dim3(256), 0, c.stream, n, ebound, mark, pos,
                       pos + n, er, eslot, where, state);
    const double *X = c.x[0].as<double>(), *Y = c.x[1].as<double>(), *Z = c.x[2].as<double>();
    hipLaunchKernelGGL(k_dt_end_prep, dim3(nb256(ebound)), dim3(256), 0, c.stream, ebound, state,
                       er, where, X, Y, Z, c.fwd.as<int>(), c.rev.as<int>(), c.status.as<int>(),
                       dparams(c));
    // Biot-Savart at every endpoint with the links before the pass
    const double *xv[3] = {X, Y, Z};
    double *op[3] = {c.bs_prim[0].as<double>(), c.bs_prim[1].as<double>(),
                     c.bs_prim[2].as<double>()};
    double *oi[3] = {c.bs_img[0].as<double>(), c.bs_img[1].as<double>(),
                     c.bs_img[2].as<double>()};
    const bool exact = !(c.params.rec_velocity == OVG_RECVEL_TREE &&
                         c.params.method == OVG_METHOD_BH);
    if (!exact) {
        // the step's own Barnes-Hut evaluator at the endpoints
        OVG_TRY(build_tree(c, xv));
        OVG_TRY(T.sub_groups.ensure(sizeof(int2) * (size_t)ebound + sizeof(int) * (size_t)ebound +
                                    64));
        OVG_TRY(T.sub_count.ensure(sizeof(int)));
        int2 *d_grp = T.sub_groups.as<int2>();
        int *d_tgt = reinterpret_cast<int *>(d_grp + ebound);
        hipLaunchKernelGGL(k_dt_groups, dim3(1), dim3(1024), 0, c.stream, ebound, state, eslot,
                           T.rank.as<int>(), T.keys[0].as<uint64_t>(),
                           list_points_per_wave(c.params), list_group_level(c.params), d_grp,
                           d_tgt, T.sub_count.as<int>());
        OVG_TRY(launch_bh_list_groups(c, xv, d_grp, d_tgt, T.sub_count.as<int>(), ebound));
    } else {
        // update_velocity(tree = NULL): the FP64 direct sum (:250-251)
        launch_prep_segments(c, xv);
        ovg_params saved = c.params;
        c.params.precision = OVG_PREC_FP64;
        const int rd = launch_direct_counted(c, xv, ebound, std::max(256, c.dt_ends_seen), eslot,
                                             state + DT_ENDS, op, oi);
        c.params = saved;
        OVG_TRY(rd);
    }
    hipLaunchKernelGGL(k_dt_end_vel, dim3(nb256(ebound)), dim3(256), 0, c.stream, ebound, state,
                       er, op[0], op[1], op[2], oi[0], oi[1], oi[2], dparams(c));
    const double cutoff = std::cos(c.params.rec_angle);
    hipLaunchKernelGGL(k_dt_pair_prep, dim3(nb256(cap)), dim3(256), 0, c.stream, cap, state, pr,
                       er, where, X, Y, Z, dparams(c), cutoff);
    const double rc = 25 * c.params.rec_dist, rc2 = rc * rc;
    if (!exact)
        hipLaunchKernelGGL(k_dt_near, dim3(ebound), dim3(256), 0, c.stream, ebound, state, er, pr,
                           X, Y, Z, dparams(c), rc2);
    GreedyArgs g{};
    g.pr = pr;
    g.er = er;
    g.where = where;
    g.fwd = c.fwd.as<int>();
    g.rev = c.rev.as<int>();
    g.X = X;
    g.Y = Y;
    g.Z = Z;
    g.chg = c.dt_chg.as<int4>();
    g.P = dparams(c);
    g.cutoff = cutoff;
    g.rc2 = rc2;
    g.exact = exact ? 1 : 0;
    g.max_chg = 2 * cap;
    g.state = state;
    hipLaunchKernelGGL(k_dt_greedy, dim3(1), dim3(64), 0, c.stream, g);
    hipLaunchKernelGGL(k_dt_clear_where, dim3(nb256(ebound)), dim3(256), 0, c.stream, ebound,
                       state, er, where);
    device_edited(c);
    return launch_status();
}

}  // namespace ovg
