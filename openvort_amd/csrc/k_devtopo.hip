// Topology passes on the device (ovg_params.topology == OVG_TOPO_DEVICE):
//   reconnect              src/vortex_dynamics.c:186-288, do_reconnection :128-175
//   eliminate_small_loops  src/tangle.c:1063-1128 (closed loops)
//   remesh                 src/tangle.c:1018-1061 with add_point :261-335,
//                          remove_point :342-355 and get_next_empty_node :201-253
// as kernels on the links, status and positions in HBM.  Nothing here waits
// for the host: the passes leave their counts in a small array of device
// counters (DT_*) that the host reads when a call needs them (sync_live).
//
// The sequential parts keep the reference's visiting order: the greedy
// reconnection and the remesh visit run per independent cluster (pairs or
// slots that no other cluster's edits can reach), one thread or wavefront
// each, with a one-wavefront kernel over the whole pass where clusters do
// not apply.  They make the same decisions as the host passes
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
    DT_VISITS,     // slots the last remesh visited
    DT_SLOW,       // endpoint velocities the last greedy pass recomposed
    DT_NEAR_OVER,  // endpoints with more rewirable pairs near them than kNear
    DT_PREM_FAIL,       // remesh by clusters: reasons it handed the pass to the sequential kernel
    DT_PREM_HINT,       // the allocator hint after its replay
    DT_PREM_NEW,        // new points of the cluster pass
    DT_PREM_CAND,       // remesh candidates of the last pass
    DT_CL_N,            // clusters of the last greedy pass by clusters
    DT_CL_MAX,          // candidate pairs in its largest cluster
    DT_PUBLIC,          // the counters ovg_topology_counters returns
    DT_DELTA_ON = DT_PUBLIC,  // direct mode: the field-change table holds every change
    DT_COUNT = 32
};
constexpr int kErrPairs = 1, kErrEnds = 2, kErrSlots = 4, kErrHeap = 8;
constexpr int kMaxPairs = 32768, kMaxEnds = 65536;
constexpr int kNear = 32;         // tree mode: rewirable pairs listed per endpoint
constexpr int kHeapLds = 4096;    // remesh: top of the heap of touched slots, in LDS
constexpr size_t kDeltaBytes = size_t(4) << 30;  // direct mode's field-change table

template <class T>
__device__ __forceinline__ T ld(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// the same at device scope: sees other lanes' atomics
template <class T>
__device__ __forceinline__ T ld_agent(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

__global__ void k_dt_iota(int n, int *__restrict__ a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
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
    __shared__ int wmin[4];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long b = __ballot(k < n && st[k] == OVG_EMPTY);
    if (lane == 0) wmin[w] = b ? k + __ffsll((long long)b) - 1 : INT_MAX;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = INT_MAX;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m = min(m, wmin[i]);
        // blocks run roughly in order: most find a lower value already there
        if (m < INT_MAX && m < ld(state + DT_LO)) atomicMin(state + DT_LO, m);
    }
}

struct RemeshArgs {
    int n;  // slots
    const int *gate;  // runs only when *gate is nonzero (null: always)
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

// The slot state the one-wavefront passes read and write, cached in LDS:
// direct-mapped by slot, write-through (every store goes to the cache and
// to global memory), so a miss always finds the current value in global
// memory.
constexpr int kCache = 2048;
struct SlotCache {
    int slot[kCache], fwd[kCache], rev[kCache], st[kCache], claim[kCache];
    double x[kCache], y[kCache], z[kCache];
};

struct RemeshCtx {
    const RemeshArgs &a;
    SlotCache &C;
    // entry of slot s (loaded on a miss)
    __device__ __forceinline__ int get(int s)
    {
        const int h = s & (kCache - 1);
        if (C.slot[h] != s) {
            const int f = ld(a.fwd + s), r = ld(a.rev + s), q = ld(a.st + s);
            const double x = ld(a.X + s), y = ld(a.Y + s), z = ld(a.Z + s);
            C.slot[h] = s;
            C.fwd[h] = f;
            C.rev[h] = r;
            C.st[h] = q;
            C.x[h] = x;
            C.y[h] = y;
            C.z[h] = z;
        }
        return h;
    }
    __device__ __forceinline__ int fwd(int s) { return C.fwd[get(s)]; }
    __device__ __forceinline__ int rev(int s) { return C.rev[get(s)]; }
    __device__ __forceinline__ int st(int s) { return C.st[get(s)]; }
    __device__ __forceinline__ void pos(int s, double &x, double &y, double &z)
    {
        const int h = get(s);
        x = C.x[h];
        y = C.y[h];
        z = C.z[h];
    }
    __device__ __forceinline__ void set_fwd(int s, int v)
    {
        C.fwd[get(s)] = v;
        a.fwd[s] = v;
    }
    __device__ __forceinline__ void set_rev(int s, int v)
    {
        C.rev[get(s)] = v;
        a.rev[s] = v;
    }
    __device__ __forceinline__ void set_st(int s, int v)
    {
        C.st[get(s)] = v;
        a.st[s] = v;
    }
    __device__ __forceinline__ void set_pos(int s, double x, double y, double z)
    {
        const int h = get(s);
        C.x[h] = x;
        C.y[h] = y;
        C.z[h] = z;
        a.X[s] = x;
        a.Y[s] = y;
        a.Z[s] = z;
    }
    // update_tangent_normal (src/tangle.c:423-502) with the current links
    __device__ void stencil(const TangleParams &P, int k, double t[3], double nr[3])
    {
        const int i1 = fwd(k), im1 = rev(k);
        const int i2 = fwd(i1), im2 = rev(im1);
        double p[5][3];
        const int sl[5] = {k, i1, i2, im1, im2};
        for (int j = 0; j < 5; ++j) pos(sl[j], p[j][0], p[j][1], p[j][2]);
        tangent_normal_pts(P, p[0][0], p[0][1], p[0][2], p[1][0], p[1][1], p[1][2], p[2][0],
                           p[2][1], p[2][2], p[3][0], p[3][1], p[3][2], p[4][0], p[4][1],
                           p[4][2], t, nr);
    }
    // Load the neighbourhoods of candidates cand[c0 .. c0 + 64) (rev^2 .. fwd^3,
    // what a visit and add_point's two stencils read) into the cache, one
    // candidate per lane.  Lanes that stage different slots for one entry
    // leave the entry to one of them whole.
    __device__ void prefetch(int c0, int n_cand, int *cbuf)
    {
        const int lane = threadIdx.x;
        const int ci = c0 + lane;
        cbuf[lane] = ci < n_cand ? a.cand[ci] : INT_MAX;
        int sl[6];
        int f[6], r[6], q[6];
        double x[6], y[6], z[6];
        for (int j = 0; j < 6; ++j) sl[j] = -1;
        if (ci < n_cand) {
            const int k = a.cand[ci];
            sl[0] = k;
            for (int j = 0; j < 6; ++j) {
                // order: k, fwd k, rev k, fwd^2 k, rev^2 k, fwd^3 k
                if (j == 1) sl[1] = f[0];
                if (j == 2) sl[2] = r[0];
                if (j == 3) sl[3] = sl[1] >= 0 ? f[1] : -1;
                if (j == 4) sl[4] = sl[2] >= 0 ? r[2] : -1;
                if (j == 5) sl[5] = sl[3] >= 0 ? f[3] : -1;
                const int s = sl[j];
                if (s < 0 || s >= a.n) {
                    sl[j] = -1;
                    f[j] = r[j] = -1;
                    continue;
                }
                f[j] = ld(a.fwd + s);
                r[j] = ld(a.rev + s);
                q[j] = ld(a.st + s);
                x[j] = ld(a.X + s);
                y[j] = ld(a.Y + s);
                z[j] = ld(a.Z + s);
            }
        }
        __syncthreads();
        for (int j = 0; j < 6; ++j)
            if (sl[j] >= 0) C.claim[sl[j] & (kCache - 1)] = sl[j];
        __syncthreads();
        for (int j = 0; j < 6; ++j) {
            const int s = sl[j];
            if (s < 0) continue;
            const int h = s & (kCache - 1);
            if (C.claim[h] != s) continue;
            C.slot[h] = s;
            C.fwd[h] = f[j];
            C.rev[h] = r[j];
            C.st[h] = q[j];
            C.x[h] = x[j];
            C.y[h] = y[j];
            C.z[h] = z[j];
        }
        __syncthreads();
    }
};

// The sequential remesh (src/tangle.c:1022-1056) over the candidate slots
// and the slots the pass's own edits reach ahead of its cursor, in slot
// order, as the host pass (ovg_remesh).  One wavefront; lanes share the
// prefetch of each 64 candidates' neighbourhoods and the allocator's scan
// for an EMPTY slot.
__global__ __launch_bounds__(64) void k_dt_remesh(RemeshArgs a)
{
    __shared__ SlotCache C;
    __shared__ int cbuf[64];  // the candidates of the current prefetch
    // a binary min-heap whose first kHeapLds entries (its top levels) live
    // in LDS and the rest in global memory
    __shared__ int lheap[kHeapLds];
    if (a.gate && *a.gate == 0) return;
    const TangleParams &P = *a.P;
    const int lane = threadIdx.x;
    const int n = a.n, n_cand = *a.n_cand;
    for (int i = lane; i < kCache; i += 64) C.slot[i] = -1;
    __syncthreads();
    RemeshCtx S{a, C};
    int hsize = 0, hint = a.state[DT_HINT], lo = a.state[DT_LO];
    int cursor = -1, ci = 0, added = 0, removed = 0, err = 0, fetched = 0, visits = 0;
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
    // known not to be EMPTY, so the scan from slot 0 starts there (global
    // memory holds every status: the cache writes through)
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
        if (ci >= fetched && ci < n_cand) {
            S.prefetch(ci, n_cand, cbuf);
            fetched = ci + 64;
        }
        const int av = ci < n_cand ? cbuf[ci - (fetched - 64)] : INT_MAX;
        const int bv = hsize ? lheap[0] : INT_MAX;
        const int k = av < bv ? av : bv;
        if (k == INT_MAX || err) break;
        if (av == k) ++ci;
        while (hsize && lheap[0] == k) pop();
        if (k <= cursor) continue;
        cursor = k;
        ++visits;
        if (k >= n || S.st(k) == OVG_EMPTY) continue;
        const int next = S.fwd(k), prev = S.rev(k);
        double x, y, z;
        S.pos(k, x, y, z);
        double lf = 0, lr = 0;
        if (next >= 0) {
            double bx, by, bz;
            S.pos(next, bx, by, bz);
            lf = dt_dist(P, x, y, z, bx, by, bz);
        }
        if (prev >= 0) {
            double bx, by, bz;
            S.pos(prev, bx, by, bz);
            lr = dt_dist(P, x, y, z, bx, by, bz);
        }
        if (next >= 0 && prev >= 0 &&
            ((lf < a.min_dist || lr < a.min_dist) && (lf + lr) < a.max_dist)) {
            // remove_point (:342-355)
            const int pv = S.rev(k), nx = S.fwd(k);
            if (pv >= 0) S.set_fwd(pv, nx);
            if (nx >= 0) S.set_rev(nx, pv);
            S.set_fwd(k, -1);
            S.set_rev(k, -1);
            S.set_st(k, OVG_EMPTY);
            if (k < lo) lo = k;
            ++removed;
            touch(k);
            touch(pv);
            touch(nx);
        }
        if (next >= 0 && lf > a.max_dist) {
            // add_point (:261-335)
            const int nxt = S.fwd(k);
            const int nw = alloc();
            if (nw < 0) {
                err |= kErrSlots;
                break;
            }
            S.set_st(nw, OVG_FREE);
            const int p = k, nn = S.fwd(p);
            double t0[3], n0[3], t1[3], n1[3];
            S.stencil(P, p, t0, n0);
            S.stencil(P, nn, t1, n1);
            double s0[3], s1[3];
            S.pos(p, s0[0], s0[1], s0[2]);
            S.pos(nn, s1[0], s1[1], s1[2]);
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
            S.set_pos(nw, np[0], np[1], np[2]);
            for (int q = 0; q < 3; ++q) {
                a.vs[q][nw] = 0.0;
                a.vels[q][nw] = 0.0;
            }
            S.set_rev(nw, p);
            S.set_fwd(nw, nn);
            S.set_fwd(p, nw);
            S.set_rev(nn, nw);
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
        a.state[DT_VISITS] = visits;
        if (err) atomicOr(a.state + DT_ERR, err);
    }
}

// union-find with path halving; the larger root goes under the smaller
__device__ __forceinline__ int uf_find(int *parent, int x)
{
    for (;;) {
        const int p = __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (p == x) return x;
        const int gp = __hip_atomic_load(parent + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (gp != p) atomicCAS(parent + x, p, gp);  // path halving
        x = p;
    }
}

__device__ __forceinline__ void uf_union(int *parent, int a, int b)
{
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        // the larger root goes under the smaller
        if (atomicCAS(parent + a, a, b) == a) return;
    }
}

// ------------------------------------------------------ remesh by clusters
// The sequential remesh made parallel without changing its result.  Visits
// interact through the links (a visit reads its point's neighbours, add_point
// two stencils: at most three links away) and through the allocator; the
// first is split off by clusters, the second replayed afterwards:
//  1. candidates within kPremChain links of each other form a cluster; one
//     thread per cluster runs the visits of its candidates, and of the slots
//     their edits reach ahead of the cursor, in slot order, on the links in
//     place; points it inserts are new points numbered in a table (links to
//     them hold -2 - id) that are never visited;
//  2. one wavefront replays get_next_empty_node over the insertions in visit
//     order: at visit c a slot is EMPTY when it was so before the pass, or a
//     visit before c removed it (a removal happens at its own slot), and no
//     earlier insertion took it;
//  3. the new points move to their slots and the links to them are
//     resolved.
// The reference would visit a new point whose slot lies ahead of the cursor
// that inserted it.  Such a visit changes nothing when the point's links
// are still those of its insertion and its remesh test fails; otherwise, as
// when two clusters reach the same slot, a cluster holds more touched slots
// than its heap or the table overflows, the links are restored from a copy
// and the one-wavefront pass (k_dt_remesh) runs instead.
constexpr int kPremChain = 6;
constexpr int kPremHeap = 32;
constexpr int kFailClash = 1, kFailHeap = 2, kFailTable = 4, kFailVisit = 8, kFailCand = 16;

struct NewPt {
    int fwd, rev;   // links; another new point j as -2 - j
    int cursor;     // the visit that inserted it
    int slot;       // its slot, from the allocator replay
    double x, y, z;
    int relinked;   // a later visit changed its links
    int pad;
};

struct PremArgs {
    int n;                     // slots
    const int *cand, *n_cand;  // candidate slots, ascending
    int cand_cap;              // candidates a cluster pass takes
    int sbits, ibits;          // key bits of a slot and of a new point's number
    int *fwd, *rev, *st;
    const double *X, *Y, *Z;
    double *vs[3], *vels[3];
    double *WX, *WY, *WZ;      // positions, written for the new points only
    const TangleParams *P;
    double min_dist, max_dist;
    int *cwhere;               // slot -> candidate index, -1
    int *parent;               // union-find over candidates
    const uint64_t *ckeys;     // candidates by (cluster, slot)
    int *owner;                // slot -> cluster that touched it, -1
    unsigned *removed;         // bitmap: removed by this pass
    unsigned *taken;           // bitmap: allocated by the replay
    unsigned *origE;           // bitmap: EMPTY before the pass
    int words;                 // 32-bit words of each bitmap
    NewPt *np;
    int np_cap;
    int *state;
};

__global__ void k_prem_where(PremArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = *a.n_cand;
    if (i == 0) {
        a.state[DT_PREM_CAND] = nc;
        if (nc > a.cand_cap) atomicOr(a.state + DT_PREM_FAIL, kFailCand);
    }
    if (i >= nc || i >= a.cand_cap) return;
    a.cwhere[a.cand[i]] = i;
    a.parent[i] = i;
}

__global__ void k_prem_union(PremArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = *a.n_cand;
    if (i >= nc || i >= a.cand_cap || a.state[DT_PREM_FAIL]) return;
    const int k = a.cand[i];
    for (int dir = 0; dir < 2; ++dir) {
        int s = k;
        for (int h = 0; h < kPremChain; ++h) {
            s = dir ? a.rev[s] : a.fwd[s];
            if (s < 0 || s == k) break;
            const int j = a.cwhere[s];
            if (j >= 0) uf_union(a.parent, i, j);
        }
    }
}

__global__ void k_prem_keys(PremArgs a, uint64_t *__restrict__ keys, int *__restrict__ vals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.cand_cap) return;
    const int nc = *a.n_cand;
    const bool on = i < nc && !a.state[DT_PREM_FAIL];
    keys[i] = on ? ((uint64_t)uf_find(a.parent, i) << a.sbits) | (uint64_t)(unsigned)a.cand[i]
                 : ~0ull;
    vals[i] = i;
}

// One cluster's visits (k_dt_remesh's, new points never visited).
__global__ void k_prem_cluster(PremArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = *a.n_cand;
    if (i >= nc || i >= a.cand_cap || a.state[DT_PREM_FAIL]) return;
    const uint64_t root = a.ckeys[i] >> a.sbits;
    if (i > 0 && (a.ckeys[i - 1] >> a.sbits) == root) return;
    const TangleParams &P = *a.P;
    const int cid = (int)root;
    const uint64_t smask = (1ull << a.sbits) - 1;
    int fail = 0;
    // the cluster's own record of the slots it reaches
    auto tag = [&](int s) {
        if (s < 0) return;
        const int o = atomicCAS(a.owner + s, -1, cid);
        if (o != -1 && o != cid) fail |= kFailClash;
    };
    auto f_of = [&](int s) { return s >= 0 ? a.fwd[s] : a.np[-2 - s].fwd; };
    auto r_of = [&](int s) { return s >= 0 ? a.rev[s] : a.np[-2 - s].rev; };
    auto set_f = [&](int s, int v) {
        if (s >= 0) {
            a.fwd[s] = v;
        } else {
            a.np[-2 - s].fwd = v;
            a.np[-2 - s].relinked = 1;
        }
    };
    auto set_r = [&](int s, int v) {
        if (s >= 0) {
            a.rev[s] = v;
        } else {
            a.np[-2 - s].rev = v;
            a.np[-2 - s].relinked = 1;
        }
    };
    auto pos = [&](int s, double &x, double &y, double &z) {
        if (s >= 0) {
            x = a.X[s];
            y = a.Y[s];
            z = a.Z[s];
        } else {
            const NewPt &q = a.np[-2 - s];
            x = q.x;
            y = q.y;
            z = q.z;
        }
    };
    auto stencil = [&](int k, double t[3], double nr[3]) {
        const int i1 = f_of(k), im1 = r_of(k);
        const int i2 = f_of(i1), im2 = r_of(im1);
        tag(i1);
        tag(im1);
        tag(i2);
        tag(im2);
        double p[5][3];
        const int sl[5] = {k, i1, i2, im1, im2};
        for (int j = 0; j < 5; ++j) pos(sl[j], p[j][0], p[j][1], p[j][2]);
        tangent_normal_pts(P, p[0][0], p[0][1], p[0][2], p[1][0], p[1][1], p[1][2], p[2][0],
                           p[2][1], p[2][2], p[3][0], p[3][1], p[3][2], p[4][0], p[4][1],
                           p[4][2], t, nr);
    };
    int heap[kPremHeap];
    int hsize = 0, cursor = -1, removed = 0;
    auto touch = [&](int sl) {
        if (sl < 0 || sl <= cursor) return;  // new points (< -1) are never visited
        if (hsize == kPremHeap) {
            fail |= kFailHeap;
            return;
        }
        int j = hsize++;
        while (j > 0 && heap[(j - 1) >> 1] > sl) {
            heap[j] = heap[(j - 1) >> 1];
            j = (j - 1) >> 1;
        }
        heap[j] = sl;
    };
    auto pop = [&]() {
        const int v = heap[--hsize];
        int j = 0;
        for (;;) {
            int ch = 2 * j + 1;
            if (ch >= hsize) break;
            if (ch + 1 < hsize && heap[ch + 1] < heap[ch]) ++ch;
            if (v <= heap[ch]) break;
            heap[j] = heap[ch];
            j = ch;
        }
        if (hsize > 0) heap[j] = v;
    };
    int ci = i;
    for (;;) {
        const int av = (ci < nc && (a.ckeys[ci] >> a.sbits) == root) ? (int)(a.ckeys[ci] & smask)
                                                                       : INT_MAX;
        const int bv = hsize ? heap[0] : INT_MAX;
        const int k = av < bv ? av : bv;
        if (k == INT_MAX || fail) break;
        if (av == k) ++ci;
        while (hsize && heap[0] == k) pop();
        if (k <= cursor) continue;
        cursor = k;
        tag(k);
        if (a.st[k] == OVG_EMPTY) continue;
        const int next = a.fwd[k], prev = a.rev[k];
        tag(next);
        tag(prev);
        double x, y, z;
        pos(k, x, y, z);
        double lf = 0, lr = 0;
        if (next != -1) {
            double bx, by, bz;
            pos(next, bx, by, bz);
            lf = dt_dist(P, x, y, z, bx, by, bz);
        }
        if (prev != -1) {
            double bx, by, bz;
            pos(prev, bx, by, bz);
            lr = dt_dist(P, x, y, z, bx, by, bz);
        }
        if (next != -1 && prev != -1 &&
            ((lf < a.min_dist || lr < a.min_dist) && (lf + lr) < a.max_dist)) {
            // remove_point (:342-355) of slot k, at visit k
            const int pv = a.rev[k], nx = a.fwd[k];
            if (pv != -1) set_f(pv, nx);
            if (nx != -1) set_r(nx, pv);
            a.fwd[k] = -1;
            a.rev[k] = -1;
            a.st[k] = OVG_EMPTY;
            atomicOr(a.removed + (k >> 5), 1u << (k & 31));
            ++removed;
            touch(k);
            touch(pv);
            touch(nx);
        }
        if (next != -1 && lf > a.max_dist) {
            // add_point (:261-335) into a numbered new point
            const int nxt = a.fwd[k];
            const int id = atomicAdd(a.state + DT_PREM_NEW, 1);
            if (id >= a.np_cap) {
                fail |= kFailTable;
                break;
            }
            const int p = k, nn = a.fwd[p];
            double t0[3], n0[3], t1[3], n1[3];
            stencil(p, t0, n0);
            stencil(nn, t1, n1);
            double s0[3], s1[3];
            pos(p, s0[0], s0[1], s0[2]);
            pos(nn, s1[0], s1[1], s1[2]);
            pwrap(P, s0[0], s0[1], s0[2], s1[0], s1[1], s1[2]);
            const double l = dist3(s0[0], s0[1], s0[2], s1[0], s1[1], s1[2]);
            double nv[3], npos[3];
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
                for (int q = 0; q < 3; ++q) npos[q] = aa[q] + bb[q];
            } else {
                for (int q = 0; q < 3; ++q) npos[q] = s0[q] + s1[q];
                for (int q = 0; q < 3; ++q) npos[q] = npos[q] * 0.5;
            }
            NewPt &q = a.np[id];
            q.fwd = nn;
            q.rev = p;
            q.cursor = k;
            q.slot = -1;
            q.x = npos[0];
            q.y = npos[1];
            q.z = npos[2];
            q.relinked = 0;
            set_f(p, -2 - id);
            set_r(nn, -2 - id);
            touch(k);
            touch(nxt);
        }
    }
    if (removed) atomicAdd(a.state + DT_REMOVED, removed);
    if (fail) atomicOr(a.state + DT_PREM_FAIL, fail);
}

// new points by the visit that inserted them: (cursor, id)
__global__ void k_prem_event_keys(PremArgs a, uint64_t *__restrict__ keys,
                                  int *__restrict__ vals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.np_cap) return;
    const int nn = min(a.state[DT_PREM_NEW], a.np_cap);
    const bool on = i < nn && !a.state[DT_PREM_FAIL];
    keys[i] = on ? ((uint64_t)(unsigned)a.np[i].cursor << a.ibits) | (unsigned)i : ~0ull;
    vals[i] = i;
}

// EMPTY before the pass, one 32-bit word of slots per thread
__global__ void k_prem_origE(int n, int words, const int *__restrict__ st,
                             unsigned *__restrict__ origE)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    unsigned b = 0;
    for (int j = 0; j < 32; ++j) {
        const int s = 32 * w + j;
        if (s < n && st[s] == OVG_EMPTY) b |= 1u << j;
    }
    origE[w] = b;
}

// get_next_empty_node (src/tangle.c:201-253) for every insertion in visit
// order; one wavefront, lanes share the scans.  At visit c a slot is EMPTY
// when it was before the pass, or a visit before c removed it (a removal
// happens at its own slot, at visit = slot), and no earlier insertion took
// it.  The scan's answer is the lower of the first slot EMPTY before the
// pass and not taken, and the first slot removed before c and not taken.
__global__ __launch_bounds__(64) void k_prem_alloc(PremArgs a, const uint64_t *__restrict__ ev)
{
    if (a.state[DT_PREM_FAIL]) return;
    const int lane = threadIdx.x;
    const int n = a.n;
    const int nn = min(a.state[DT_PREM_NEW], a.np_cap);
    int hint = a.state[DT_HINT];
    auto bit = [&](const unsigned *b, int s) {
        return (ld_agent(b + (s >> 5)) >> (s & 31)) & 1u;
    };
    auto empty_at = [&](int s, int c) {
        if (s < 0 || s >= n || bit(a.taken, s)) return false;
        return bit(a.origE, s) || (bit(a.removed, s) && s < c);
    };
    // first slot >= from set in `bits` and not taken (n when none); 64
    // words (2048 slots) per round
    auto first = [&](const unsigned *bits, int from) {
        for (int w0 = from >> 5; w0 < a.words; w0 += 64) {
            const int w = w0 + lane;
            unsigned v = 0;
            if (w < a.words) {
                v = ld_agent(bits + w) & ~ld_agent(a.taken + w);
                if (w == (from >> 5)) v &= ~0u << (from & 31);
            }
            const unsigned long long m = __ballot(v != 0);
            if (m) {
                const int l = __ffsll((long long)m) - 1;
                const unsigned vw = __shfl(v, l, 64);
                const int sl = 32 * (w0 + l) + __ffs(vw) - 1;
                return sl < n ? sl : n;
            }
        }
        return n;
    };
    int po = first(a.origE, 0), pr = first(a.removed, 0);
    for (int e = 0; e < nn; ++e) {
        const uint64_t key = ev[e];
        const int id = (int)(key & ((1ull << a.ibits) - 1)), c = (int)(key >> a.ibits);
        int got = -1;
        if (hint > 0 && hint < n && empty_at(hint, c)) {
            got = hint;
        } else {
            if (po < n && bit(a.taken, po)) po = first(a.origE, po);
            if (pr < n && bit(a.taken, pr)) pr = first(a.removed, pr);
            got = po;
            if (pr < c && pr < got) got = pr;
            if (got >= n) got = -1;
        }
        if (got < 0) {
            if (lane == 0) atomicOr(a.state + DT_ERR, kErrSlots);
            if (lane == 0) atomicOr(a.state + DT_PREM_FAIL, kFailTable);
            return;
        }
        hint = empty_at(got + 1, c) ? got + 1 : -1;
        if (lane == 0) {
            atomicOr(a.taken + (got >> 5), 1u << (got & 31));
            a.np[id].slot = got;
        }
        __syncthreads();
    }
    if (lane == 0) a.state[DT_PREM_HINT] = hint;
}

__device__ __forceinline__ int prem_slot(const PremArgs &a, int v)
{
    return v <= -2 ? a.np[-2 - v].slot : v;
}

// The reference visits a new point whose slot is ahead of its insertion's
// cursor: that visit must change nothing.
__global__ void k_prem_check(PremArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nn = min(a.state[DT_PREM_NEW], a.np_cap);
    if (i >= nn || a.state[DT_PREM_FAIL]) return;
    const NewPt &q = a.np[i];
    if (q.slot <= q.cursor) return;
    bool acts = q.relinked != 0;
    if (!acts) {
        const TangleParams &P = *a.P;
        double bx, by, bz, cx, cy, cz;
        auto pos = [&](int s, double &x, double &y, double &z) {
            if (s >= 0) {
                x = a.X[s]; y = a.Y[s]; z = a.Z[s];
            } else {
                const NewPt &w = a.np[-2 - s];
                x = w.x; y = w.y; z = w.z;
            }
        };
        pos(q.fwd, bx, by, bz);
        pos(q.rev, cx, cy, cz);
        const double lf = dt_dist(P, q.x, q.y, q.z, bx, by, bz);
        const double lr = dt_dist(P, q.x, q.y, q.z, cx, cy, cz);
        acts = ((lf < a.min_dist || lr < a.min_dist) && (lf + lr) < a.max_dist) ||
               lf > a.max_dist;
    }
    if (acts) atomicOr(a.state + DT_PREM_FAIL, kFailVisit);
}

__global__ void k_prem_apply(PremArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nn = min(a.state[DT_PREM_NEW], a.np_cap);
    if (i == 0 && !a.state[DT_PREM_FAIL]) {
        a.state[DT_ADDED] = nn;
        a.state[DT_HINT] = a.state[DT_PREM_HINT];
    }
    if (i >= nn || a.state[DT_PREM_FAIL]) return;
    const NewPt &q = a.np[i];
    const int s = q.slot;
    a.WX[s] = q.x;
    a.WY[s] = q.y;
    a.WZ[s] = q.z;
    for (int d = 0; d < 3; ++d) {
        a.vs[d][s] = 0.0;
        a.vels[d][s] = 0.0;
    }
    a.st[s] = OVG_FREE;
    const int f = prem_slot(a, q.fwd), r = prem_slot(a, q.rev);
    a.fwd[s] = f;
    a.rev[s] = r;
    if (q.fwd >= 0) a.rev[q.fwd] = s;
    if (q.rev >= 0) a.fwd[q.rev] = s;
}

// fallback: the links and status before the pass
__global__ void k_prem_restore(int n, const int *__restrict__ state, int *__restrict__ fwd,
                               int *__restrict__ rev, int *__restrict__ st,
                               const int *__restrict__ sf, const int *__restrict__ sr,
                               const int *__restrict__ ss)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !state[DT_PREM_FAIL]) return;
    fwd[i] = sf[i];
    rev[i] = sr[i];
    st[i] = ss[i];
}

__global__ void k_prem_unwhere(PremArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < *a.n_cand && i < a.cand_cap) a.cwhere[a.cand[i]] = -1;
}

__global__ void k_prem_copy(int n, const int *__restrict__ a, int *__restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
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

__global__ void k_dt_pair_prep(int cap, const int *__restrict__ state, PairRec *__restrict__ pr,
                               const EndRec *__restrict__ er, const int *__restrict__ where,
                               const double *__restrict__ X, const double *__restrict__ Y,
                               const double *__restrict__ Z, const TangleParams *__restrict__ Pp,
                               double cutoff, double4 *__restrict__ sphere)
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
    // a sphere about x_k holding the four points where a segment this pair
    // rewires can start (k, l and their reverse neighbours), for the near lists
    double rad2 = 0;
    const int segs[3] = {l, kr, lr};
    for (int q = 0; q < 3; ++q) {
        const int g = segs[q];
        rad2 = fmax(rad2, dt_near2(P, X[g], Y[g], Z[g], xk, yk, zk));
    }
    sphere[j] = make_double4(xk, yk, zk, sqrt(rad2));
}


// Tree mode: for every endpoint, the pairs whose rewiring could change a
// segment within the correction radius of it (a rewired segment starts at
// k, l or their reverse neighbours), ascending; more than kNear: -1.  Pairs
// are screened by a sphere about x_k that holds those four points.
__global__ void k_dt_near(int bound, int *__restrict__ state, EndRec *__restrict__ er,
                          const PairRec *__restrict__ pr, const double4 *__restrict__ sphere,
                          const double *__restrict__ X, const double *__restrict__ Y,
                          const double *__restrict__ Z, const TangleParams *__restrict__ Pp,
                          double rc)
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
    const double lim = rc * rc * (1 + 1e-9);
    for (int j = threadIdx.x; j < np; j += blockDim.x) {
        const double4 b = sphere[j];
        const double reach = rc * (1 + 1e-9) + b.w * (1 + 1e-12);
        if (dt_near2(P, b.x, b.y, b.z, px, py, pz) >= reach * reach) continue;
        const PairRec &R = pr[j];
        const int segs[4] = {R.k, R.l, er[R.ek].rev, er[R.el].rev};
        bool hit = false;
        for (int q = 0; q < 4 && !hit; ++q) {
            const int g = segs[q];
            hit = dt_near2(P, X[g], Y[g], Z[g], px, py, pz) < lim;
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
            atomicAdd(state + DT_NEAR_OVER, 1);
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

// Target groups of the endpoints for any endpoint count, as the host path
// (launch_bh_list_targets) forms them: the endpoints in tree order (the
// sorted points flagged as endpoints, compacted by a scan), cut where two
// consecutive endpoints lie in different octree cells of the group level
// and after every points_per_wave endpoints of a cell's run.
__global__ void k_dt_rank_flag(int n, const int *__restrict__ spts, const int *__restrict__ where,
                               int *__restrict__ f)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) f[r] = where[spts[r]] >= 0;
}

__global__ void k_dt_rank_compact(int n, const int *__restrict__ f, const int *__restrict__ pos,
                                  const int *__restrict__ spts, const uint64_t *__restrict__ keys,
                                  int shift, int *__restrict__ tgt, uint64_t *__restrict__ cid)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n || !f[r]) return;
    const int i = pos[r];
    tgt[i] = spts[r];
    // shift < 0: a group level beyond the keys' levels, every endpoint a
    // cell of its own
    cid[i] = shift < 0 ? (uint64_t)i : keys[r] >> shift;
}

// first index of the run of cell c in the ascending cid[0, m), and one past
// its last
__device__ __forceinline__ int run_lo(const uint64_t *cid, int m, uint64_t c)
{
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cid[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int run_hi(const uint64_t *cid, int m, uint64_t c)
{
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cid[mid] <= c) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_dt_group_flag(int bound, const int *__restrict__ m_p,
                                const uint64_t *__restrict__ cid, int ppw, int *__restrict__ gf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bound) return;
    const int m = min(*m_p, bound);
    gf[i] = i < m && (i - run_lo(cid, m, cid[i])) % ppw == 0;
}

__global__ void k_dt_group_emit(int bound, const int *__restrict__ m_p,
                                const uint64_t *__restrict__ cid, const int *__restrict__ gf,
                                const int *__restrict__ gpos, int ppw, int2 *__restrict__ grp,
                                int *__restrict__ n_groups)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *n_groups = gpos[bound];
    if (i >= bound || !gf[i]) return;
    const int m = min(*m_p, bound);
    grp[gpos[i]] = make_int2(i, min(ppw, run_hi(cid, m, cid[i]) - i));
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
    const double *delta;  // direct mode: field change of pair change r at endpoint e,
                          // 3 doubles at (r * ends + e) * 3 (state[DT_DELTA_ON])
    double *bsr;          // with the table: each endpoint's Biot-Savart with every
                          // rewiring so far, component q at q * ends + e
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

// Direct mode: every rewiring changes every endpoint's velocity, by the
// field of its two new segments minus that of the two old ones.  The
// changes a pair would make are fixed by the links before the pass
// (do_reconnection's case and the endpoints' recorded links), so the table
// of each candidate change's field change at each endpoint is computed up
// front, in parallel, with the arithmetic the sequential pass uses
// (dt_segment_all, new minus old); the pass then only adds table entries in
// rewiring order.  Row 2 j + s is pair j's change s (s = 0: the segment
// from k, or from l; s = 1: the other), as the pass numbers its changes.
// Tables that would exceed `cap` entries are not filled (DT_DELTA_ON = 0)
// and the pass evaluates the changes itself.  With the table, the pass
// keeps every endpoint's running sum (bsr, from the pre-pass values) and
// adds each rewiring's two rows to all of them as it happens, so an
// endpoint's sum at any point is the one the sequential pass forms, with
// the additions in the same order.
__global__ void k_dt_delta(GreedyArgs a, size_t cap, double *__restrict__ delta)
{
    const int np = a.state[DT_PAIRS], ne = a.state[DT_ENDS];
    const size_t total = 2 * (size_t)np * (size_t)ne;
    const bool on = total > 0 && total <= cap;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.state[DT_DELTA_ON] = on;
    if (!on) return;
    const TangleParams &P = *a.P;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x)
        for (int q = 0; q < 3; ++q) a.bsr[(size_t)q * ne + e] = a.er[e].bs[q];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / (size_t)ne), e = (int)(i - (size_t)r * ne);
        const PairRec &R = a.pr[r >> 1];
        const int kind = R.info & 3;
        if (!kind) continue;  // never rewired: the row is never read
        const EndRec &K = a.er[R.ek], &L = a.er[R.el];
        int m, of, nf;
        if (kind == 1) {
            m = (r & 1) ? L.rev : R.k;
            of = (r & 1) ? R.l : K.fwd;
            nf = (r & 1) ? K.fwd : R.l;
        } else {
            m = (r & 1) ? K.rev : R.l;
            of = (r & 1) ? R.k : L.fwd;
            nf = (r & 1) ? L.fwd : R.k;
        }
        const int p = a.er[e].slot;
        double vo[3] = {0, 0, 0}, vn[3] = {0, 0, 0};
        dt_segment_all(a, P, m, of, p, vo);
        dt_segment_all(a, P, m, nf, p, vn);
        for (int q = 0; q < 3; ++q) delta[3 * i + q] = vn[q] - vo[q];
    }
}

// The sequential greedy pass of src/vortex_dynamics.c:224-284 in (k, l)
// order, as host_recon.h greedy_pass: one wavefront.  Pairs are taken in
// batches of 64 whose records, and their endpoints' values from before the
// pass, are staged in LDS together; a pair whose endpoints no rewiring has
// reached since is decided from those alone.
__global__ __launch_bounds__(64) void k_dt_greedy(GreedyArgs a)
{
    __shared__ unsigned fbits[kGreedyWords], sbits[kGreedyWords], rbits[kMaxPairs / 32];
    __shared__ double sd[64][3];
    __shared__ int sfound[64];
    __shared__ int bk[64], bl[64], bek[64], bel[64], binfo[64], bnk[64], bnl[64];
    __shared__ int bnear[64][2][kNear];
    __shared__ double bdx[64][3], bvk[64][3], bvl[64][3];
    const TangleParams &P = *a.P;
    const int lane = threadIdx.x;
    // tree mode with every near list whole: the cluster kernel takes the pass
    if (!a.exact && a.state[DT_NEAR_OVER] == 0) return;
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
    for (int i = lane; i < kMaxPairs / 32; i += 64) rbits[i] = 0;
    __syncthreads();
    auto bit = [&](const unsigned *b, int e) { return e >= 0 && ((b[e >> 5] >> (e & 31)) & 1); };
    auto set = [&](unsigned *b, int e) {
        if (e >= 0) b[e >> 5] |= 1u << (e & 31);
    };
    int nch = 0, nrec = 0, slow = 0;
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
    // (the table, when filled, holds the same differences: k_dt_delta)
    const bool use_delta = a.exact && a.state[DT_DELTA_ON];
    const int n_ends = a.state[DT_ENDS];
    auto apply = [&](int cnt, int e, int p, double bs[3]) {
        for (int b0 = 0; b0 < cnt; b0 += 64) {
            const int nb = min(64, cnt - b0);
            __syncthreads();
            if (lane < nb) {
                const int c = sfound[b0 + lane];
                if (use_delta) {
                    const int r = 2 * ld(&a.chg[c].w) + (c & 1);
                    const double *d = a.delta + 3 * ((size_t)r * n_ends + e);
                    for (int q = 0; q < 3; ++q) sd[lane][q] = d[q];
                } else {
                    const int4 ch = a.chg[c];
                    double vo[3] = {0, 0, 0}, vn[3] = {0, 0, 0};
                    dt_segment_all(a, P, ch.x, ch.y, p, vo);
                    dt_segment_all(a, P, ch.x, ch.z, p, vn);
                    for (int q = 0; q < 3; ++q) sd[lane][q] = vn[q] - vo[q];
                }
            }
            __syncthreads();
            for (int i = 0; i < nb; ++i)
                for (int q = 0; q < 3; ++q) bs[q] += sd[i][q];
        }
    };
    // velocity of endpoint e; vel0 / near_n / near: its values before the pass
    auto velocity = [&](int e, const double t[3], const double nr[3], const double *vel0,
                        int near_n, const int *near, double v[3]) {
        const bool st_e = bit(sbits, e);
        if (!st_e) {
            bool clean;
            if (a.exact) {
                clean = nch == 0;
            } else {
                clean = near_n >= 0;
                for (int i = 0; i < near_n && clean; ++i) clean = !bit(rbits, near[i]);
            }
            if (clean) {
                for (int q = 0; q < 3; ++q) v[q] = vel0[q];
                return;
            }
        }
        ++slow;
        EndRec &E = a.er[e];
        const int p = E.slot;
        double bs[3] = {ld(&E.bs[0]), ld(&E.bs[1]), ld(&E.bs[2])};
        bool found = false;
        if (use_delta) {
            // the running sum, kept up to date at every rewiring
            found = nch > 0;
            for (int q = 0; q < 3; ++q) bs[q] = ld_agent(a.bsr + (size_t)q * n_ends + e);
        } else if (a.exact) {
            // every rewiring so far, added once into the endpoint's running sum
            found = nch > 0;
            int app = ld(&E.applied);
            while (app < nch) {
                const int nb = min(64, nch - app);
                __syncthreads();
                if (lane < nb) sfound[lane] = app + lane;
                __syncthreads();
                apply(nb, e, p, bs);
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
            if (near_n >= 0) {
                for (int i = 0; i < near_n; ++i) {
                    if (!bit(rbits, near[i])) continue;
                    const int c0 = ld(&a.pr[near[i]].chg);
                    for (int c = c0; c < c0 + 2; ++c) {
                        const int g = ld(&a.chg[c].x);
                        if (dt_near2(P, a.X[g], a.Y[g], a.Z[g], px, py, pz) < a.rc2) {
                            if (lane == 0) sfound[cnt] = c;
                            ++cnt;
                        }
                    }
                }
                __syncthreads();
                apply(cnt, e, p, bs);
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
                    apply(nb, e, p, bs);
                    cnt += nb;
                }
            }
            found = cnt > 0;
        }
        if (!found && !st_e) {
            for (int q = 0; q < 3; ++q) v[q] = vel0[q];
            return;
        }
        dt_compose(P, E.f_lia, t, nr, bs, v);
    };

    int cur_k = -1;
    bool skip = false, have_t = false, have_v = false;
    double tk[3], nk[3], vk[3];
    for (int b0 = 0; b0 < np; b0 += 64) {
        const int nbat = min(64, np - b0);
        __syncthreads();
        if (lane < nbat) {
            const PairRec &R = a.pr[b0 + lane];
            const int ek = R.ek, el = R.el;
            bk[lane] = R.k;
            bl[lane] = R.l;
            bek[lane] = ek;
            bel[lane] = el;
            binfo[lane] = R.info;
            const EndRec &K = a.er[ek], &L = a.er[el];
            bnk[lane] = K.near_n;
            bnl[lane] = L.near_n;
            for (int q = 0; q < 3; ++q) {
                bdx[lane][q] = R.dx[q];
                bvk[lane][q] = K.vel[q];
                bvl[lane][q] = L.vel[q];
            }
            if (!a.exact)
                for (int i = 0; i < kNear; ++i) {
                    bnear[lane][0][i] = K.near[i];
                    bnear[lane][1][i] = L.near[i];
                }
        }
        __syncthreads();
        for (int jb = 0; jb < nbat; ++jb) {
            const int j = b0 + jb;
            const int k = bk[jb], l = bl[jb], ek = bek[jb], el = bel[jb];
            if (k != cur_k) {
                cur_k = k;
                skip = bit(fbits, ek);
                have_t = false;
                have_v = false;
            }
            if (skip || bit(fbits, el)) continue;
            const int info = binfo[jb];
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
                velocity(ek, tk, nk, bvk[jb], bnk[jb], bnear[jb][0], vk);
                have_v = true;
            }
            if (!have_tl) tangent(el, tl, nl);
            double vl[3];
            velocity(el, tl, nl, bvl[jb], bnl[jb], bnear[jb][1], vl);
            double dv[3];
            for (int q = 0; q < 3; ++q) dv[q] = vk[q] - vl[q];
            if (dot3(bdx[jb][0], bdx[jb][1], bdx[jb][2], dv[0], dv[1], dv[2]) > 0) continue;
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
            a.pr[j].chg = nch;
            set(rbits, j);
            nch += 2;
            ++nrec;
            if (use_delta) {
                // both changes into every endpoint's running sum, in order
                const double *d0 = a.delta + 3 * (size_t)(2 * j) * n_ends;
                const double *d1 = d0 + 3 * (size_t)n_ends;
                for (int e = lane; e < n_ends; e += 64)
                    for (int q = 0; q < 3; ++q) {
                        double *b = a.bsr + (size_t)q * n_ends + e;
                        *b = (ld_agent(b) + d0[3 * e + q]) + d1[3 * e + q];
                    }
                __syncthreads();
            }
            // flags: the pair, its old neighbours (and the new ones, which are
            // among them)
            set(fbits, ek);
            set(fbits, el);
            set(fbits, K.wf);
            set(fbits, K.wr);
            set(fbits, L.wf);
            set(fbits, L.wr);
            // stale stencils: the six slots and their current neighbours,
            // one slot per lane
            {
                const int six[6] = {k, l, kf, kr, lf, lr};
                int w0 = -1, w1 = -1, w2 = -1;
                if (lane < 6) {
                    const int sl = six[lane];
                    const int f = ld(a.fwd + sl), r = ld(a.rev + sl);
                    w0 = a.where[sl];
                    w1 = f >= 0 ? a.where[f] : -1;
                    w2 = r >= 0 ? a.where[r] : -1;
                }
                __syncthreads();
                for (int q = 0; q < 6; ++q) {
                    set(sbits, __shfl(w0, q, 64));
                    set(sbits, __shfl(w1, q, 64));
                    set(sbits, __shfl(w2, q, 64));
                }
                __syncthreads();
            }
            skip = true;
        }
    }
    if (lane == 0) {
        a.state[DT_REC] = nrec;
        a.state[DT_CHANGES] = nch;
        a.state[DT_SLOW] = slow;
    }
}

// ---- the greedy pass by clusters (tree-mode velocities) -----------------
// A rewiring reaches later pairs only through the flags and stale stencils
// it sets (slots within two links of its endpoints), the links it writes
// (within one link) and, in tree mode, the field corrections within the
// correction radius (the endpoints' near lists).  Joining each pair with
// every endpoint within two links of its own endpoints and with the pairs
// on its endpoints' near lists splits the pairs into clusters that no
// other cluster's rewiring reaches; each is processed alone in (k, l)
// order, one thread per cluster, which gives the sequential pass's result.
// Union-find nodes: pairs [0, cap), endpoints [cap, cap + ebound).

__global__ void k_cl_pairs(int cap, const int *__restrict__ state, const PairRec *__restrict__ pr,
                           const EndRec *__restrict__ er, const int *__restrict__ where,
                           const int *__restrict__ fwd, const int *__restrict__ rev,
                           int *__restrict__ parent)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cap || j >= state[DT_PAIRS] || state[DT_ENDS] == 0) return;
    const PairRec &R = pr[j];
    uf_union(parent, j, cap + R.ek);
    uf_union(parent, j, cap + R.el);
    for (int side = 0; side < 2; ++side) {
        const EndRec &E = er[side ? R.el : R.ek];
        const int f = E.fwd, r = E.rev;
        const int sl[4] = {f, r, f >= 0 ? fwd[f] : -1, r >= 0 ? rev[r] : -1};
        for (int q = 0; q < 4; ++q)
            if (sl[q] >= 0) {
                const int e = where[sl[q]];
                if (e >= 0) uf_union(parent, j, cap + e);
            }
    }
}

__global__ void k_cl_near(int cap, int bound, const int *__restrict__ state,
                          const EndRec *__restrict__ er, int *__restrict__ parent)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= bound || e >= state[DT_ENDS] || state[DT_NEAR_OVER]) return;
    const EndRec &E = er[e];
    for (int i = 0; i < E.near_n; ++i) uf_union(parent, cap + e, E.near[i]);
}

// pair j -> (cluster root << 15 | j), sortable; the rest sort last
__global__ void k_cl_keys(int cap, const int *__restrict__ state, int *__restrict__ parent,
                          uint64_t *__restrict__ keys, int *__restrict__ vals)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cap) return;
    const bool on = j < state[DT_PAIRS] && state[DT_ENDS] > 0;
    keys[j] = on ? ((uint64_t)uf_find(parent, j) << 15) | (uint64_t)j : ~0ull;
    vals[j] = j;
}

// cluster starts: the first pair of each cluster, in any order
__global__ void k_cl_starts(int cap, const uint64_t *__restrict__ keys, int *__restrict__ state,
                            int *__restrict__ starts)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int np = state[DT_PAIRS];
    if (i >= cap || i >= np || state[DT_ENDS] == 0 || state[DT_NEAR_OVER]) return;
    if (i > 0 && (keys[i - 1] >> 15) == (keys[i] >> 15)) return;
    starts[atomicAdd(state + DT_CL_N, 1)] = i;
}

struct ClusterArgs {
    int cap;
    const uint64_t *keys;   // pairs by (cluster, j)
    const int *starts;      // first pair of each cluster (state[DT_CL_N] of them)
    PairRec *pr;
    EndRec *er;
    const int *where;
    int *fwd, *rev;
    const double *X, *Y, *Z;
    int4 *chg;               // pair j's changes at 2 j, 2 j + 1
    int *flag, *stale;       // per endpoint
    const TangleParams *P;
    double cutoff, rc2;
    int *state;
};

// host_recon.h greedy_pass restricted to one cluster, one wavefront per
// cluster.  The lanes run the pass in step (every lane holds the same
// values and makes the same writes; what the pass wrote is read back with
// ld); a field correction evaluates its two segments' 27 image terms on
// separate lanes and adds them in image order, as the sequential sum does.
__global__ __launch_bounds__(64) void k_cl_greedy(ClusterArgs a)
{
    __shared__ double sv[2][32][3];
    __shared__ int sok[2][32];
    const int lane = threadIdx.x;
    const int np = a.state[DT_PAIRS];
    const int ncl = a.state[DT_CL_N];
    if (a.state[DT_ENDS] == 0 || a.state[DT_NEAR_OVER]) return;
    const TangleParams &P = *a.P;
    auto tangent = [&](int e, double t[3], double nr[3]) {
        const EndRec &E = a.er[e];
        if (!ld(a.stale + e)) {
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
    // bs += field of (m -> nf) - field of (m -> of) at slot p, each summed
    // over the images in order (primary box skipped when p is a segment end)
    auto correct = [&](int m, int of, int nf, int p, double bs[3]) {
        const int half = lane >> 5, im = (lane & 31) - 1;
        const int f = half ? nf : of;
        double v[3] = {0, 0, 0};
        int ok = 0;
        if (im < P.n_images && !(im < 0 && (m == p || f == p))) {
            const double xm = a.X[m], ym = a.Y[m], zm = a.Z[m];
            double bx = a.X[f], by = a.Y[f], bz = a.Z[f];
            pwrap(P, xm, ym, zm, bx, by, bz);
            const double rx = a.X[p], ry = a.Y[p], rz = a.Z[p];
            double qx = rx, qy = ry, qz = rz;
            if (im >= 0) image_point(P, im, rx, ry, rz, qx, qy, qz);
            ok = seg_field64(xm - qx, ym - qy, zm - qz, bx - qx, by - qy, bz - qz, P.f_bs, v[0],
                             v[1], v[2]);
        }
        __syncthreads();
        for (int q = 0; q < 3; ++q) sv[half][lane & 31][q] = v[q];
        sok[half][lane & 31] = ok;
        __syncthreads();
        double vo[3] = {0, 0, 0}, vn[3] = {0, 0, 0};
        for (int k = 0; k <= P.n_images; ++k) {
            if (sok[0][k])
                for (int q = 0; q < 3; ++q) vo[q] += sv[0][k][q];
            if (sok[1][k])
                for (int q = 0; q < 3; ++q) vn[q] += sv[1][k][q];
        }
        for (int q = 0; q < 3; ++q) bs[q] += vn[q] - vo[q];
    };
    for (int ci = blockIdx.x; ci < ncl; ci += gridDim.x) {
        const int i = a.starts[ci];
        const uint64_t root = a.keys[i] >> 15;
        int nrec = 0;
        auto velocity = [&](int e, const double t[3], const double nr[3], double v[3]) {
            const EndRec &E = a.er[e];
            const int p = E.slot;
            double bs[3] = {E.bs[0], E.bs[1], E.bs[2]};
            bool found = false;
            const double px = a.X[p], py = a.Y[p], pz = a.Z[p];
            // corrections come from this cluster's rewirings only
            const int nq = nrec > 0 ? E.near_n : 0;
            for (int q = 0; q < nq; ++q) {
                const int c0 = ld(&a.pr[E.near[q]].chg);
                if (c0 < 0) continue;
                for (int c = c0; c < c0 + 2; ++c) {
                    const int m = ld(&a.chg[c].x), of = ld(&a.chg[c].y), nf = ld(&a.chg[c].z);
                    if (!(dt_near2(P, a.X[m], a.Y[m], a.Z[m], px, py, pz) < a.rc2)) continue;
                    correct(m, of, nf, p, bs);
                    found = true;
                }
            }
            if (!found && !ld(a.stale + e)) {
                for (int q = 0; q < 3; ++q) v[q] = E.vel[q];
                return;
            }
            dt_compose(P, E.f_lia, t, nr, bs, v);
        };
        int cur_k = -1, npairs = 0;
        bool skip = false, have_t = false, have_v = false;
        double tk[3], nk[3], vk[3];
        for (int t = i; t < np && (a.keys[t] >> 15) == root; ++t) {
            ++npairs;
            const int j = (int)(a.keys[t] & 0x7fff);
            PairRec &R = a.pr[j];
            const int k = R.k, l = R.l, ek = R.ek, el = R.el;
            if (k != cur_k) {
                cur_k = k;
                skip = ld(a.flag + ek);
                have_t = false;
                have_v = false;
            }
            if (skip || ld(a.flag + el)) continue;
            const int info = R.info;
            double tl[3], nl[3];
            bool have_tl = false;
            if (!ld(a.stale + ek) && !ld(a.stale + el)) {
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
            const EndRec &K = a.er[ek], &L = a.er[el];
            const int kf = K.fwd, kr = K.rev, lf = L.fwd, lr = L.rev;
            if (kind == 1) {
                a.fwd[k] = l;
                a.rev[l] = k;
                a.rev[kf] = lr;
                a.fwd[lr] = kf;
                a.chg[2 * j] = make_int4(k, kf, l, j);
                a.chg[2 * j + 1] = make_int4(lr, l, kf, j);
            } else {
                a.fwd[l] = k;
                a.rev[k] = l;
                a.rev[lf] = kr;
                a.fwd[kr] = lf;
                a.chg[2 * j] = make_int4(l, lf, k, j);
                a.chg[2 * j + 1] = make_int4(kr, k, lf, j);
            }
            R.chg = 2 * j;
            ++nrec;
            const int fl[6] = {ek, el, K.wf, K.wr, L.wf, L.wr};
            for (int q = 0; q < 6; ++q)
                if (fl[q] >= 0) a.flag[fl[q]] = 1;
            const int six[6] = {k, l, kf, kr, lf, lr};
            for (int q = 0; q < 6; ++q) {
                const int sl = six[q];
                const int w[3] = {a.where[sl], a.where[ld(a.fwd + sl)], a.where[ld(a.rev + sl)]};
                for (int d = 0; d < 3; ++d)
                    if (w[d] >= 0) a.stale[w[d]] = 1;
            }
            skip = true;
        }
        if (lane == 0) {
            if (nrec) atomicAdd(a.state + DT_REC, nrec);
            atomicMax(a.state + DT_CL_MAX, npairs);
        }
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
    c.prem_seen = std::max(c.prem_seen, st[DT_PREM_CAND]);
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
    int st[DT_COUNT];
    OVG_TRY(devtopo_counters(c, st));
    out[0] = st[DT_REC];
    out[1] = st[DT_LOOPS];
    out[2] = st[DT_ADDED];
    out[3] = st[DT_REMOVED];
    return OVG_OK;
}

int devtopo_counters(Ctx &c, int out[22])
{
    static_assert(DT_PUBLIC == 22, "ovg_topology_counters returns 22 counters");
    OVG_TRY(dt_state_ready(c));
    OVG_TRY(sync_live(c));
    OVG_HIP(hipMemcpyAsync(out, c.dt_state.p, sizeof(int) * DT_PUBLIC, hipMemcpyDeviceToHost,
                           c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
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

// The cluster pass (remesh by clusters, above); leaves DT_PREM_FAIL nonzero
// with the links and status as before the pass when it hands the pass back.
static int prem_queue(Ctx &c, const int *cand, const int *n_cand, double min_dist,
                      double max_dist)
{
    int *state = c.dt_state.as<int>();
    const int n = c.n;
    int cap = 16384;
    while (cap < (1 << 22) && cap < 2 * c.prem_seen) cap *= 2;
    const int np_cap = cap;
    const int words = (n + 31) / 32 + 1;
    OVG_TRY(c.dt_where.ensure(sizeof(int) * (size_t)(n + 1)));
    OVG_TRY(c.prem_buf.ensure(sizeof(int) * ((size_t)cap + 4 * (size_t)n + 3 * (size_t)words + 64) +
                              sizeof(NewPt) * (size_t)np_cap + 256));
    int *parent = c.prem_buf.as<int>();
    int *owner = parent + cap;
    int *snap = owner + n;  // fwd, rev, status
    unsigned *removed = reinterpret_cast<unsigned *>(snap + 3 * (size_t)n);
    unsigned *taken = removed + words;
    unsigned *origE = taken + words;
    NewPt *np = reinterpret_cast<NewPt *>(
        (reinterpret_cast<uintptr_t>(origE + words) + 63) & ~(uintptr_t)63);
    TreeBufs &T = c.tree;
    for (int b = 0; b < 2; ++b) {
        OVG_TRY(T.keys[b].ensure(sizeof(uint64_t) * (size_t)cap));
        OVG_TRY(T.vals[b].ensure(sizeof(int) * (size_t)cap));
    }
    ++c.tree_id;
    hipLaunchKernelGGL(k_dt_fill, dim3(nb256(n)), dim3(256), 0, c.stream, n,
                       c.dt_where.as<int>(), -1);
    for (int d : {DT_PREM_FAIL, DT_PREM_NEW, DT_ADDED, DT_REMOVED})
        hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, d, 0);
    const int *fw = c.fwd.as<int>(), *rv = c.rev.as<int>(), *ss = c.status.as<int>();
    hipLaunchKernelGGL(k_prem_copy, dim3(nb256(n)), dim3(256), 0, c.stream, n, fw, snap);
    hipLaunchKernelGGL(k_prem_copy, dim3(nb256(n)), dim3(256), 0, c.stream, n, rv, snap + n);
    hipLaunchKernelGGL(k_prem_copy, dim3(nb256(n)), dim3(256), 0, c.stream, n, ss,
                       snap + 2 * (size_t)n);
    hipLaunchKernelGGL(k_dt_fill, dim3(nb256(n)), dim3(256), 0, c.stream, n, owner, -1);
    hipLaunchKernelGGL(k_dt_fill, dim3(nb256(2 * words)), dim3(256), 0, c.stream, 2 * words,
                       reinterpret_cast<int *>(removed), 0);
    hipLaunchKernelGGL(k_prem_origE, dim3(nb256(words)), dim3(256), 0, c.stream, n, words, ss,
                       origE);
    PremArgs a{};
    a.n = n;
    a.cand = cand;
    a.n_cand = n_cand;
    a.cand_cap = cap;
    a.sbits = 1;
    while ((1LL << a.sbits) < (long long)n) ++a.sbits;
    a.ibits = 1;
    while ((1 << a.ibits) < cap) ++a.ibits;  // candidates and new points: both below cap
    a.fwd = c.fwd.as<int>();
    a.rev = c.rev.as<int>();
    a.st = c.status.as<int>();
    a.X = c.x[0].as<double>();
    a.Y = c.x[1].as<double>();
    a.Z = c.x[2].as<double>();
    a.WX = c.x[0].as<double>();
    a.WY = c.x[1].as<double>();
    a.WZ = c.x[2].as<double>();
    for (int q = 0; q < 3; ++q) {
        a.vs[q] = c.vs[q].as<double>();
        a.vels[q] = c.vels[q].as<double>();
    }
    a.P = dparams(c);
    a.min_dist = min_dist;
    a.max_dist = max_dist;
    a.cwhere = c.dt_where.as<int>();
    a.parent = parent;
    a.ckeys = T.keys[0].as<uint64_t>();
    a.owner = owner;
    a.removed = removed;
    a.taken = taken;
    a.origE = origE;
    a.words = words;
    a.np = np;
    a.np_cap = np_cap;
    a.state = state;
    hipLaunchKernelGGL(k_prem_where, dim3(nb256(cap)), dim3(256), 0, c.stream, a);
    hipLaunchKernelGGL(k_prem_union, dim3(nb256(cap)), dim3(256), 0, c.stream, a);
    hipLaunchKernelGGL(k_prem_keys, dim3(nb256(cap)), dim3(256), 0, c.stream, a,
                       T.keys[0].as<uint64_t>(), T.vals[0].as<int>());
    // (cluster, slot): the root is a candidate index below cap
    int passes = (a.sbits + a.ibits + 7) / 8;
    if (passes & 1) ++passes;
    OVG_TRY(radix_sort(c, cap, 0, 8 * (passes - 1)));
    hipLaunchKernelGGL(k_prem_cluster, dim3(nb256(cap)), dim3(256), 0, c.stream, a);
    hipLaunchKernelGGL(k_prem_event_keys, dim3(nb256(np_cap)), dim3(256), 0, c.stream, a,
                       T.keys[0].as<uint64_t>(), T.vals[0].as<int>());
    // (cursor, number): the cursor is a slot
    OVG_TRY(radix_sort(c, np_cap, 0, 8 * (passes - 1)));
    hipLaunchKernelGGL(k_prem_alloc, dim3(1), dim3(64), 0, c.stream, a,
                       T.keys[0].as<const uint64_t>());
    hipLaunchKernelGGL(k_prem_check, dim3(nb256(np_cap)), dim3(256), 0, c.stream, a);
    hipLaunchKernelGGL(k_prem_apply, dim3(nb256(np_cap)), dim3(256), 0, c.stream, a);
    hipLaunchKernelGGL(k_prem_restore, dim3(nb256(n)), dim3(256), 0, c.stream, n, state,
                       c.fwd.as<int>(), c.rev.as<int>(), c.status.as<int>(), snap, snap + n,
                       snap + 2 * (size_t)n);
    hipLaunchKernelGGL(k_prem_unwhere, dim3(nb256(cap)), dim3(256), 0, c.stream, a);
    return OVG_OK;
}

int devtopo_remesh(Ctx &c, double min_dist, double max_dist)
{
    OVG_TRY(dt_state_ready(c));
    int *state = c.dt_state.as<int>();
    // room for the insertions: the reference doubles the slots when none is
    // EMPTY (get_next_empty_node, src/tangle.c:241-253); doubling earlier
    // leaves every allocation the same slot, since the scan finds the old
    // slots first either way.  With fewer free slots than a quarter of the
    // points (the live count here may be an over-estimate), the pass's
    // insertion bound is read once, and the slots double only when it may
    // not fit.
    const int free_slots = c.n - c.n_live;
    const int *cand = nullptr, *n_cand = nullptr;
    if (c.n > 0 && free_slots < std::max(4096, c.n_live / 4)) {
        OVG_TRY(positions_change(c));
        OVG_TRY(c.rc_room.ensure(sizeof(unsigned long long)));
        hipLaunchKernelGGL(k_dt_fill, dim3(1), dim3(64), 0, c.stream, 2, c.rc_room.as<int>(), 0);
        OVG_TRY(remesh_candidates_queue(c, min_dist, max_dist, &cand, &n_cand,
                                        c.rc_room.as<unsigned long long>()));
        unsigned long long room = 0;
        OVG_HIP(hipMemcpyAsync(&room, c.rc_room.p, sizeof(room), hipMemcpyDeviceToHost,
                               c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
        if ((unsigned long long)std::max(free_slots, 0) < room + 1024) {
            OVG_TRY(sync_mirrors(c));
            long long grown = 2 * (long long)c.n;
            while (grown - c.n_live < (long long)room + 1024) grown *= 2;
            if (grown > INT32_MAX / 2)
                return set_error(OVG_ELIMIT, "device remesh: %llu possible insertions exceed "
                                             "the slot limit", room);
            OVG_TRY(grow_capacity(c, (int)grown));
            cand = n_cand = nullptr;  // scanned again over the grown slots
        }
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
    if (!cand) OVG_TRY(remesh_candidates_queue(c, min_dist, max_dist, &cand, &n_cand));
    OVG_TRY(prem_queue(c, cand, n_cand, min_dist, max_dist));
    // the one-wavefront pass, for the passes the clusters hand back
    hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_LO, n);
    hipLaunchKernelGGL(k_dt_first_empty, dim3(nb256(n)), dim3(256), 0, c.stream, n,
                       c.status.as<int>(), state);
    // each visit touches at most 6 slots ahead of the cursor, and every
    // slot is visited at most once
    const size_t gcap = 6 * (size_t)n + 64;
    OVG_TRY(c.dt_chg.ensure(std::max(c.dt_chg.bytes, sizeof(int) * gcap)));
    RemeshArgs a{};
    a.n = n;
    a.gate = state + DT_PREM_FAIL;
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
    hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_NEAR_OVER, 0);
    hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_CHANGES, 0);
    hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_SLOW, 0);
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
    OVG_TRY(c.dt_pairs.ensure((sizeof(PairRec) + sizeof(double4)) * (size_t)cap));
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
    int ex_pairs = 0, ex_ends = 0;  // direct mode: the pass's counts
    if (!exact) {
        // the step's own Barnes-Hut evaluator at the endpoints
        OVG_TRY(build_tree(c, xv));
        OVG_TRY(T.sub_groups.ensure(sizeof(int2) * (size_t)ebound + sizeof(int) * (size_t)ebound +
                                    64));
        OVG_TRY(T.sub_count.ensure(sizeof(int)));
        int2 *d_grp = T.sub_groups.as<int2>();
        int *d_tgt = reinterpret_cast<int *>(d_grp + ebound);
        // the endpoints in tree order and their target groups
        const int nl = c.n_live;
        const size_t s1 = ((size_t)nl + 64) & ~(size_t)63, s2 = ((size_t)ebound + 64) & ~(size_t)63;
        OVG_TRY(c.dt_grp.ensure(sizeof(int) * (2 * s1 + 2 * s2) + sizeof(uint64_t) * s2));
        int *gfl = c.dt_grp.as<int>(), *gpo = gfl + s1, *ef = gpo + s1, *epo = ef + s2;
        uint64_t *cid = reinterpret_cast<uint64_t *>(epo + s2);
        hipLaunchKernelGGL(k_dt_rank_flag, dim3(nb256(nl)), dim3(256), 0, c.stream, nl,
                           T.sorted_pts.as<int>(), where, gfl);
        OVG_TRY(launch_exclusive_scan(c, gfl, gpo, nl, gpo + nl, T.scan_tmp));
        hipLaunchKernelGGL(k_dt_rank_compact, dim3(nb256(nl)), dim3(256), 0, c.stream, nl, gfl,
                           gpo, T.sorted_pts.as<int>(), T.keys[0].as<uint64_t>(),
                           list_group_level(c.params) > 16
                               ? -1
                               : tuned_key_cell_shift(list_group_level(c.params)),
                           d_tgt, cid);
        const int ppw = list_points_per_wave(c.params);
        hipLaunchKernelGGL(k_dt_group_flag, dim3(nb256(ebound)), dim3(256), 0, c.stream, ebound,
                           gpo + nl, cid, ppw, ef);
        OVG_TRY(launch_exclusive_scan(c, ef, epo, ebound, epo + ebound, T.scan_tmp));
        hipLaunchKernelGGL(k_dt_group_emit, dim3(nb256(ebound)), dim3(256), 0, c.stream, ebound,
                           gpo + nl, cid, ef, epo, ppw, d_grp, T.sub_count.as<int>());
        OVG_TRY(launch_bh_list_groups(c, xv, d_grp, d_tgt, T.sub_count.as<int>(), ebound));
    } else {
        // update_velocity(tree = NULL): the FP64 direct sum (:250-251).
        // This mode waits once for the pass's pair and endpoint counts, so
        // the direct sum is planned for the endpoints there are (enough
        // source chunks to fill the device) and the field-change table is
        // sized to them.
        int cnt[DT_COUNT];
        OVG_HIP(hipMemcpyAsync(cnt, state, sizeof(cnt), hipMemcpyDeviceToHost, c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
        ex_pairs = std::max(0, std::min(cnt[DT_PAIRS], cap));
        ex_ends = std::max(0, std::min(cnt[DT_ENDS], ebound));
        launch_prep_segments(c, xv);
        ovg_params saved = c.params;
        c.params.precision = OVG_PREC_FP64;
        const int rd = launch_direct(c, xv, ex_ends, eslot, op, oi, 0);
        c.params = saved;
        OVG_TRY(rd);
    }
    hipLaunchKernelGGL(k_dt_end_vel, dim3(nb256(ebound)), dim3(256), 0, c.stream, ebound, state,
                       er, op[0], op[1], op[2], oi[0], oi[1], oi[2], dparams(c));
    const double cutoff = std::cos(c.params.rec_angle);
    double4 *sphere = reinterpret_cast<double4 *>(pr + cap);
    hipLaunchKernelGGL(k_dt_pair_prep, dim3(nb256(cap)), dim3(256), 0, c.stream, cap, state, pr,
                       er, where, X, Y, Z, dparams(c), cutoff, sphere);
    const double rc = 25 * c.params.rec_dist, rc2 = rc * rc;
    if (!exact)
        hipLaunchKernelGGL(k_dt_near, dim3(ebound), dim3(256), 0, c.stream, ebound, state, er, pr,
                           sphere, X, Y, Z, dparams(c), rc);
    if (!exact) {
        // clusters of pairs that no other cluster's rewiring reaches
        const int nodes = cap + ebound;
        OVG_TRY(c.dt_cl.ensure(sizeof(int) * ((size_t)nodes + 2 * (size_t)ebound + cap) + 64));
        int *parent = c.dt_cl.as<int>();
        int *flag = parent + nodes;
        int *stale = flag + ebound;
        int *starts = stale + ebound;
        hipLaunchKernelGGL(k_dt_iota, dim3(nb256(nodes)), dim3(256), 0, c.stream, nodes, parent);
        hipLaunchKernelGGL(k_dt_fill, dim3(nb256(2 * ebound)), dim3(256), 0, c.stream,
                           2 * ebound, flag, 0);
        hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_CL_N, 0);
        hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(64), 0, c.stream, state, DT_CL_MAX, 0);
        hipLaunchKernelGGL(k_cl_pairs, dim3(nb256(cap)), dim3(256), 0, c.stream, cap, state, pr,
                           er, where, c.fwd.as<int>(), c.rev.as<int>(), parent);
        hipLaunchKernelGGL(k_cl_near, dim3(nb256(ebound)), dim3(256), 0, c.stream, cap, ebound,
                           state, er, parent);
        hipLaunchKernelGGL(k_cl_keys, dim3(nb256(cap)), dim3(256), 0, c.stream, cap, state,
                           parent, T.keys[0].as<uint64_t>(), T.vals[0].as<int>());
        ++c.tree_id;
        OVG_TRY(radix_sort(c, cap, 0, 24));
        hipLaunchKernelGGL(k_cl_starts, dim3(nb256(cap)), dim3(256), 0, c.stream, cap,
                           T.keys[0].as<uint64_t>(), state, starts);
        ClusterArgs ca{};
        ca.cap = cap;
        ca.keys = T.keys[0].as<uint64_t>();
        ca.starts = starts;
        ca.pr = pr;
        ca.er = er;
        ca.where = where;
        ca.fwd = c.fwd.as<int>();
        ca.rev = c.rev.as<int>();
        ca.X = X;
        ca.Y = Y;
        ca.Z = Z;
        ca.chg = c.dt_chg.as<int4>();
        ca.flag = flag;
        ca.stale = stale;
        ca.P = dparams(c);
        ca.cutoff = cutoff;
        ca.rc2 = rc2;
        ca.state = state;
        hipLaunchKernelGGL(k_cl_greedy, dim3(std::min(cap, 16 * c.n_cu)), dim3(64), 0, c.stream,
                           ca);
    }
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
    if (exact) {
        // the table of field changes, when it fits in kDeltaBytes
        size_t entries = 2 * (size_t)ex_pairs * (size_t)ex_ends;
        if (entries * 3 * sizeof(double) > kDeltaBytes ||
            c.dt_delta.ensure((entries + ex_ends) * 3 * sizeof(double) + 64) != OVG_OK) {
            clear_error();
            entries = 0;
        }
        g.delta = c.dt_delta.as<double>();
        g.bsr = c.dt_delta.p ? c.dt_delta.as<double>() + 3 * entries : nullptr;
        hipLaunchKernelGGL(k_dt_delta, dim3(8 * c.n_cu), dim3(256), 0, c.stream, g, entries,
                           c.dt_delta.as<double>());
    }
    hipLaunchKernelGGL(k_dt_greedy, dim3(1), dim3(64), 0, c.stream, g);
    hipLaunchKernelGGL(k_dt_clear_where, dim3(nb256(ebound)), dim3(256), 0, c.stream, ebound,
                       state, er, where);
    device_edited(c);
    return launch_status();
}

}  // namespace ovg
