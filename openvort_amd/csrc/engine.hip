// Host side of the C ABI: context, state upload/download, the parameter
// block and the sequencing of the hot path (update_tangle, rk4_step,
// enforce_periodic_boundaries).  Kernels live in k_*.hip.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <queue>
#include <vector>

#include "engine_internal.h"
#include "host_phys.h"

struct ovg_ctx {
    ovg::Ctx c;
};

namespace ovg {

// Device bytes held by every DBuf of the process, and the cap set by
// ovg_set_device_memory_limit (0 = none).
static std::atomic<size_t> g_dev_in_use{0};
static std::atomic<size_t> g_dev_limit{0};
// set while this thread captures a context's stream into a graph: buffers
// must not move under the captured work (a grown buffer is freed and
// reallocated)
static thread_local bool t_capturing = false;
static thread_local bool t_capture_failed = false;  // a call in the capture failed

int DBuf::ensure(size_t n)
{
    if (n <= bytes) return OVG_OK;
    if (t_capturing) {
        t_capture_failed = true;
        return set_error(OVG_EINVAL, "a device buffer would grow during graph capture: run the "
                         "captured calls once before capturing them");
    }
    if (p) {
        hipFree(p);
        g_dev_in_use -= bytes;
        p = nullptr;
        bytes = 0;
    }
    size_t want = n < 256 ? 256 : n;
    const size_t cap = g_dev_limit.load();
    const size_t held = g_dev_in_use.fetch_add(want) + want;
    if (cap && held > cap) {
        g_dev_in_use -= want;
        return set_error(OVG_ENOMEM, "device buffer of %zu bytes exceeds the memory limit "
                         "(%zu of %zu bytes held)", want, held - want, cap);
    }
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        g_dev_in_use -= want;
        p = nullptr;
        return set_error(OVG_ENOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    bytes = want;
    return OVG_OK;
}

void DBuf::release()
{
    if (p) {
        hipFree(p);
        g_dev_in_use -= bytes;
    }
    p = nullptr;
    bytes = 0;
}

// ---- small stream-ordered helper kernels ---------------------------------
__global__ void k_set_params(TangleParams p, TangleParams *dst) { *dst = p; }

__global__ void k_aos_to_soa(int n, const double *__restrict__ aos, double *__restrict__ x,
                             double *__restrict__ y, double *__restrict__ z)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = aos[3 * i];
    y[i] = aos[3 * i + 1];
    z[i] = aos[3 * i + 2];
}

__global__ void k_soa_to_aos(int n, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, double *__restrict__ aos)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aos[3 * i] = x[i];
    aos[3 * i + 1] = y[i];
    aos[3 * i + 2] = z[i];
}

__global__ void k_split_pairs(int n, const int *__restrict__ pairs, int *__restrict__ a,
                              int *__restrict__ b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a[i] = pairs[2 * i];
    b[i] = pairs[2 * i + 1];
}

__global__ void k_join_pairs(int n, const int *__restrict__ a, const int *__restrict__ b,
                             int *__restrict__ pairs)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pairs[2 * i] = a[i];
    pairs[2 * i + 1] = b[i];
}

__global__ void k_fill_zero(int n, double *__restrict__ a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

static inline int grid_for(int n, int bs = 256) { return (n + bs - 1) / bs; }

void launch_soa_to_aos(Ctx &c, const double *x, const double *y, const double *z, double *aos,
                       hipStream_t stream)
{
    if (c.n == 0) return;
    hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(c.n)), dim3(256), 0, stream, c.n, x, y, z,
                       aos);
}

void launch_soa_to_aos(Ctx &c, const double *x, const double *y, const double *z, double *aos)
{
    launch_soa_to_aos(c, x, y, z, aos, c.stream);
}

void launch_split_pairs(Ctx &c, const int *pairs)
{
    if (c.n == 0) return;
    hipLaunchKernelGGL(k_split_pairs, dim3(grid_for(c.n)), dim3(256), 0, c.stream, c.n, pairs,
                       c.fwd.as<int>(), c.rev.as<int>());
}

// reference get_v for one flow at time t (src/external_velocity.c:13-81)
static void flow_at(const ovg_flow &f, double t, double out[3])
{
    out[0] = out[1] = out[2] = 0;
    if (f.type == OVG_FLOW_NONE) return;
    double str = f.strength;
    if (f.type == OVG_FLOW_DECREASING) {
        if (t < f.time_of_linear_decrease)
            str = f.starting_strength -
                  t * (f.starting_strength - f.strength) / f.time_of_linear_decrease;
        else
            str = f.strength;
    }
    double d[3] = {f.direction[0], f.direction[1], f.direction[2]};
    if (!(d[0] == 0 && d[1] == 0 && d[2] == 0)) {
        double l = 0;
        for (int k = 0; k < 3; ++k) l += d[k] * d[k];
        l = std::sqrt(l);
        for (int k = 0; k < 3; ++k) d[k] = d[k] * (1 / l);
    }
    for (int k = 0; k < 3; ++k) out[k] = d[k] * str;
}

TangleParams host_tparams(const ovg_params &p, double t)
{
    TangleParams P;
    std::memset(&P, 0, sizeof P);
    for (int k = 0; k < 3; ++k) {
        P.lo[k] = p.box_lo[k];
        P.hi[k] = p.box_hi[k];
        P.L[k] = p.box_hi[k] - p.box_lo[k];
        P.halfL[k] = P.L[k] / 2;
    }
    P.periodic[0] = p.wall[0] == OVG_WALL_PERIODIC;
    P.periodic[1] = p.wall[2] == OVG_WALL_PERIODIC;
    P.periodic[2] = p.wall[4] == OVG_WALL_PERIODIC;
    for (int f = 0; f < 6; ++f) P.wrap_face[f] = p.wall[f] == OVG_WALL_PERIODIC;
    P.n_images = p.n_images;
    for (int j = 0; j < p.n_images; ++j) {
        for (int k = 0; k < 3; ++k) P.Ls[j][k] = P.L[k] * p.images[j][k];
        for (int f = 0; f < 6; ++f)
            if (p.image_reflect[j] >> f & 1) {
                const int a = f / 2;  // LEFT/RIGHT x, BACK/FRONT y, DOWN/UP z
                P.refl[j] |= 1 << a;
                P.W2[j][a] = 2 * (f % 2 ? p.box_hi[a] : p.box_lo[a]);
            }
    }
    P.f_bs = p.kappa / 4 / M_PI;
    P.kappa = p.kappa;
    P.a0 = p.vortex_width;
    P.sqrtE = std::sqrt(M_E);
    flow_at(p.vs, t, P.evs);
    flow_at(p.vn, t, P.evn);
    flow_at(p.vb, t, P.evb);
    P.alpha = p.alpha;
    P.alpha_p = p.alpha_p;
    P.use_mf = p.use_mutual_friction;
    P.theta = p.theta;
    P.crit_scale = 1.0 / (p.theta * p.theta);
    P.leaf_size = p.leaf_size;
    P.group_level = list_group_level(p);
    P.near2 = p.mixed_near * p.mixed_near;
    P.n_tangles = 1;
    P.tid = nullptr;
    return P;
}

static bool flow_depends_on_time(const ovg_params &p)
{
    return p.vn.type == OVG_FLOW_DECREASING || p.vs.type == OVG_FLOW_DECREASING ||
           p.vb.type == OVG_FLOW_DECREASING;
}

// A forest's parameter table is written once, when it is uploaded (its
// flows do not depend on time); a single tangle's entry at every call.
void push_params(Ctx &c, double t)
{
    if (c.n_tangles > 1) {
        forest_table(c, false);  // an error shows in the caller's launch_status
        return;
    }
    TangleParams P = host_tparams(c.params, t);
    hipLaunchKernelGGL(k_set_params, dim3(1), dim3(1), 0, c.stream, P,
                       c.tparams.as<TangleParams>());
}

// Calls that take one tangle only: OVG_EUNSUPPORTED on a forest.
static int single_tangle(const Ctx &c, const char *what)
{
    if (c.n_tangles > 1)
        return set_error(OVG_EUNSUPPORTED,
                         "%s: this context holds a forest of %d tangles (ovg_upload_forest); "
                         "it runs the evaluation and the RK4 step only",
                         what, c.n_tangles);
    return OVG_OK;
}

static int check_ready(const ovg_ctx *ctx, bool need_state = true)
{
    if (!ctx) return set_error(OVG_EINVAL, "context is NULL");
    if (!ctx->c.have_params) return set_error(OVG_EINVAL, "ovg_set_params has not been called");
    if (need_state && ctx->c.n == 0) return set_error(OVG_EINVAL, "no tangle uploaded");
    return OVG_OK;
}

static double *dp(DBuf &b) { return b.as<double>(); }

struct Triple {
    double *v[3];
};
static Triple tri(DBuf *b)
{
    return Triple{{b[0].as<double>(), b[1].as<double>(), b[2].as<double>()}};
}

static hipEvent_t take_event(Ctx &c)
{
    hipEvent_t e = nullptr;
    if (!c.ev_pool.empty()) {
        e = c.ev_pool.back();
        c.ev_pool.pop_back();
    } else {
        hipEventCreate(&e);
    }
    return e;
}

static void mark(Ctx &c, hipEvent_t *slot)
{
    if (!c.timing) return;
    *slot = take_event(c);
    hipEventRecord(*slot, c.stream);
}

// Sum the pending (start, tree, Biot-Savart) event triples into c.stats.
static void resolve_events(Ctx &c)
{
    if (c.ev_pending.empty()) return;
    hipStreamSynchronize(c.stream);
    for (size_t i = 0; i + 2 < c.ev_pending.size(); i += 3) {
        float a = 0, b = 0;
        hipEventElapsedTime(&a, c.ev_pending[i], c.ev_pending[i + 1]);
        hipEventElapsedTime(&b, c.ev_pending[i + 1], c.ev_pending[i + 2]);
        c.stats.ms_tree += a;
        c.stats.ms_bs += b;
        c.stats.bs_launches += 1;
    }
    for (auto e : c.ev_pending) c.ev_pool.push_back(e);
    c.ev_pending.clear();
}

// Biot-Savart part of one evaluation at positions x: writes bs_prim,
// bs_img by slot.
// phase: kEvalPlain, kEvalStep (update_tangle: a step's first evaluation,
// which records its interaction lists when reuse_lists is set) or
// kEvalStage (an RK4 stage of that step, which replays them).
int eval_bs(Ctx &c, const double *const x[3], int phase)
{
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    mark(c, &e0);
    if (c.params.method == OVG_METHOD_BH) {
        const bool lists = c.params.traverse == OVG_TRAVERSE_LIST && c.params.reuse_lists;
        if (lists && phase == kEvalStage && c.rec_ok && c.rec_tree == c.tree_id &&
            c.rec_n == c.n_live) {
            OVG_TRY(refit_tree(c, x));
            c.list_mode = kReplay;
        } else {
            OVG_TRY(build_tree(c, x));
            c.rec_ok = false;
            c.list_mode = lists && phase == kEvalStep ? kRecord : kWalk;
        }
        mark(c, &e1);
        const int r = launch_bh(c, x);
        if (r == OVG_OK && c.list_mode == kRecord) {
            c.rec_ok = true;
            c.rec_tree = c.tree_id;
            c.rec_n = c.n_live;
        }
        c.list_mode = kWalk;
        OVG_TRY(r);
    } else {
        launch_prep_segments(c, x);
        mark(c, &e1);
        Triple bp = tri(c.bs_prim), bi = tri(c.bs_img);
        OVG_TRY(launch_direct(c, x, c.n_live, c.live.as<int>(), bp.v, bi.v, 0));
        // every (target, segment, image) pair except the two LIA segments
        double pairs = (double)c.n_live * c.n_src * (1 + c.params.n_images) - 2.0 * c.n_live;
        c.stats.interactions += pairs;
        if (c.params.precision == OVG_PREC_MIXED) c.stats.interactions_fp32 += pairs;
    }
    mark(c, &e2);
    if (c.timing) {
        c.ev_pending.push_back(e0);
        c.ev_pending.push_back(e1);
        c.ev_pending.push_back(e2);
    }
    return OVG_OK;
}

// Counters accumulate from one ovg_get_stats call to the next.
static int fetch_stats(Ctx &c, ovg_stats *out)
{
    resolve_events(c);
    unsigned long long cnt[2] = {0, 0};
    int tc[4] = {0, 0, 0, 0};
    OVG_HIP(hipMemcpyAsync(cnt, c.counter64.p, 16, hipMemcpyDeviceToHost, c.stream));
    if (c.tree.counters.p)
        OVG_HIP(hipMemcpyAsync(tc, c.tree.counters.p, 16, hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    *out = c.stats;
    out->interactions += (double)cnt[0] + (double)cnt[1];
    out->interactions_fp32 += (double)cnt[1];
    out->tree_nodes = tc[1];
    out->n_points = c.n_live;
    std::memset(&c.stats, 0, sizeof c.stats);
    OVG_HIP(hipMemsetAsync(c.counter64.p, 0, 16, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    return OVG_OK;
}

}  // namespace ovg

using namespace ovg;

extern "C" size_t ovg_set_device_memory_limit(size_t bytes) { return g_dev_limit.exchange(bytes); }

extern "C" size_t ovg_device_memory_in_use(void) { return g_dev_in_use.load(); }

extern "C" int ovg_create(ovg_ctx **out, int device)
{
    clear_error();
    if (!out) return set_error(OVG_EINVAL, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return set_error(OVG_EDEVICE, "no HIP device available");
    if (device < 0 || device >= ndev)
        return set_error(OVG_EINVAL, "device %d out of range (have %d)", device, ndev);
    hipDeviceProp_t prop;
    OVG_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(OVG_EDEVICE, "device %d is %s; this library is built for gfx950", device,
                         prop.gcnArchName);
    OVG_HIP(hipSetDevice(device));
    ovg_ctx *ctx = new ovg_ctx();
    Ctx &c = ctx->c;
    c.device = device;
    c.n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return set_error(OVG_EDEVICE, "hipStreamCreate failed");
    }
    for (auto &e : c.ev) hipEventCreate(&e);
    // position downloads overlap the context stream's work (default priority)
    if (hipStreamCreateWithFlags(&c.copy_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c.xpin_from, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c.xpin_done, hipEventDisableTiming) != hipSuccess) {
        ovg_destroy(ctx);
        return set_error(OVG_EDEVICE, "copy stream / events could not be created");
    }
    int r;
    if ((r = c.tparams.ensure(sizeof(TangleParams))) || (r = c.flags.ensure(64)) ||
        (r = c.counter64.ensure(64))) {
        ovg_destroy(ctx);
        return r;
    }
    if (hipMemsetAsync(c.flags.p, 0, 64, c.stream) != hipSuccess ||
        hipMemsetAsync(c.counter64.p, 0, 64, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        ovg_destroy(ctx);
        return set_error(OVG_EDEVICE, "clearing the context's device flags failed");
    }
    // the allocation-free entry point's scratch, once per device
    if ((r = direct_scratch_init(device)) != OVG_OK) {
        ovg_destroy(ctx);
        return r;
    }
    *out = ctx;
    return launch_status();
}

extern "C" int ovg_destroy(ovg_ctx *ctx)
{
    if (!ctx) return OVG_OK;
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    if (c.stream) hipStreamSynchronize(c.stream);
    if (c.copy_stream) hipStreamSynchronize(c.copy_stream);
    DBuf *all[] = {&c.fwd, &c.rev, &c.status, &c.pinw, &c.live, &c.src, &c.partial, &c.dpart, &c.seg_ids,
                   &c.tparams, &c.flags, &c.counter64, &c.rc_cell, &c.rc_sorted, &c.rc_start,
                   &c.rc_pairs, &c.rc_count, &c.rc_tmp, &c.sm_buf, &c.split_part,
                   &c.loop_off, &c.loop_slots, &c.lt_buf, &c.sm_fft, &c.sm_big, &c.sm_tw,
                   &c.sm_gstate, &c.topo_stage, &c.xaos, &c.rc_room, &c.loop_open,
                   &c.dt_state, &c.dt_where, &c.dt_mark, &c.dt_pairs, &c.dt_ends, &c.dt_chg,
                   &c.dt_cl, &c.dt_delta, &c.prem_buf, &c.dt_grp};
    for (auto *b : all) b->release();
    for (int q = 0; q < 3; ++q) {
        c.x[q].release(); c.vels[q].release(); c.vs[q].release(); c.tan[q].release();
        c.nor[q].release(); c.move[q].release(); c.k[q].release(); c.stan[q].release();
        c.snor[q].release(); c.bs_prim[q].release(); c.bs_img[q].release();
        c.xs[0][q].release(); c.xs[1][q].release();
    }
    for (auto &b : c.seg) b.release();
    TreeBufs &T = c.tree;
    DBuf *tb[] = {&T.keys[0], &T.keys[1], &T.vals[0], &T.vals[1], &T.hist, &T.delta,
                  &T.counts, &T.base, &T.nodes, &T.meta, &T.level_hist, &T.level_list,
                  &T.scan_tmp, &T.counters, &T.sorted_pts, &T.sx, &T.sy, &T.sz, &T.sqx,
                  &T.sqy, &T.sqz, &T.grp_a, &T.grp_b, &T.grp_c, &T.groups,
                  &T.rank, &T.sub_tgt, &T.sub_groups, &T.sub_count, &T.rec,
                  &T.rec_count, &T.fix, &T.fix_tiles, &T.roots};
    for (auto *b : tb) b->release();
    c.tid.release();
    if (c.graph_exec) hipGraphExecDestroy(c.graph_exec);
    c.graph_exec = nullptr;
    if (c.h_xpin) hipHostFree(c.h_xpin);
    c.h_xpin = nullptr;
    for (auto &e : c.ev)
        if (e) hipEventDestroy(e);
    for (auto e : c.ev_pending) hipEventDestroy(e);
    for (auto e : c.ev_pool) hipEventDestroy(e);
    for (auto e : c.markers)
        if (e) hipEventDestroy(e);
    if (c.sm_hmeta_done) hipEventDestroy(c.sm_hmeta_done);
    if (c.xpin_from) hipEventDestroy(c.xpin_from);
    if (c.xpin_done) hipEventDestroy(c.xpin_done);
    if (c.copy_stream) hipStreamDestroy(c.copy_stream);
    if (c.stream) hipStreamDestroy(c.stream);
    delete ctx;
    return OVG_OK;
}

static int validate_params(const ovg_params *p)
{
    for (int k = 0; k < 3; ++k)
        if (!(p->box_hi[k] > p->box_lo[k]) || !std::isfinite(p->box_hi[k] - p->box_lo[k]))
            return set_error(OVG_EINVAL, "box axis %d is empty or not finite", k);
    for (int f = 0; f < 6; ++f)
        if (p->wall[f] != OVG_WALL_OPEN && p->wall[f] != OVG_WALL_PERIODIC &&
            p->wall[f] != OVG_WALL_MIRROR)
            return set_error(OVG_EINVAL, "wall[%d] = %d is not a wall kind", f, p->wall[f]);
    if (p->n_images < 0 || p->n_images > kMaxImages)
        return set_error(OVG_EINVAL, "n_images = %d out of range", p->n_images);
    bool mirrored = false;
    for (int j = 0; j < p->n_images; ++j) {
        const int m = p->image_reflect[j];
        if (m & ~63)
            return set_error(OVG_EINVAL, "image_reflect[%d] = %d is not a face mask", j, m);
        for (int a = 0; a < 3; ++a)
            if ((m >> (2 * a) & 3) == 3)
                return set_error(OVG_EINVAL,
                                 "image_reflect[%d] mirrors through both faces of axis %d", j, a);
        mirrored |= m != 0;
    }
    if (mirrored && p->reuse_lists)
        return set_error(OVG_EUNSUPPORTED,
                         "mirror images: reuse_lists (recorded interaction lists) takes shift "
                         "images only");
    if (p->method != OVG_METHOD_DIRECT && p->method != OVG_METHOD_BH)
        return set_error(OVG_EINVAL, "unknown method %d", p->method);
    if (p->precision != OVG_PREC_FP64 && p->precision != OVG_PREC_MIXED)
        return set_error(OVG_EINVAL, "unknown precision %d", p->precision);
    if (p->method == OVG_METHOD_BH && !(p->theta > 0))
        return set_error(OVG_EINVAL, "theta must be > 0");
    if (p->leaf_size < 1 || p->leaf_size > 64)
        return set_error(OVG_EINVAL, "leaf_size must be in [1, 64]");
    if (p->traverse != OVG_TRAVERSE_POINT && p->traverse != OVG_TRAVERSE_GROUP &&
        p->traverse != OVG_TRAVERSE_LIST)
        return set_error(OVG_EINVAL, "unknown traversal %d", p->traverse);
    if (p->points_per_wave != 0 && p->points_per_wave != 8 && p->points_per_wave != 16 &&
        p->points_per_wave != 21 && p->points_per_wave != 32 && p->points_per_wave != 64)
        return set_error(OVG_EINVAL, "points_per_wave must be 0, 8, 16, 21, 32 or 64");
    if (p->group_level < 0 || p->group_level > 21)
        return set_error(OVG_EINVAL, "group_level must be in [0, 21]");
    if (p->reuse_lists != 0 && p->reuse_lists != 1)
        return set_error(OVG_EINVAL, "reuse_lists must be 0 or 1");
    if (p->list_stack != 0 && (p->list_stack < 128 || p->list_stack > 384 || p->list_stack % 64))
        return set_error(OVG_EINVAL, "list_stack must be 0 or 128..384 in steps of 64");
    if (p->list_tile != 0 && p->list_tile != 96 && p->list_tile != 128)
        return set_error(OVG_EINVAL, "list_tile must be 0, 96 or 128");
    if (p->list_blocks < 0 || p->list_blocks > 64)
        return set_error(OVG_EINVAL, "list_blocks must be in [0, 64]");
    if (p->pin_mode != 0 && p->pin_mode != OVG_PINNED && p->pin_mode != OVG_PINNED_SLIP)
        return set_error(OVG_EINVAL, "pin_mode must be 0, OVG_PINNED or OVG_PINNED_SLIP");
    if (p->traverse == OVG_TRAVERSE_POINT &&
        (p->leaf_size != 1 || p->precision != OVG_PREC_FP64) && p->method == OVG_METHOD_BH)
        return set_error(OVG_EINVAL,
                         "point traversal is the reference's FP64 split-until-one tree: "
                         "leaf_size 1 and FP64 only");
    if (!(p->kappa > 0) || !(p->vortex_width > 0))
        return set_error(OVG_EINVAL, "kappa and vortex_width must be > 0");
    return OVG_OK;
}

extern "C" int ovg_set_params(ovg_ctx *ctx, const ovg_params *p)
{
    clear_error();
    if (!ctx || !p) return set_error(OVG_EINVAL, "NULL argument");
    OVG_TRY(validate_params(p));
    Ctx &c = ctx->c;
    OVG_TRY(single_tangle(c, "ovg_set_params"));
    hipSetDevice(c.device);
    c.params = *p;
    c.rec_ok = false;  // recorded interaction lists belong to the old settings
    c.have_params = true;
    push_params(c, 0.0);
    return launch_status();
}

extern "C" int ovg_get_params(const ovg_ctx *ctx, ovg_params *p)
{
    if (!ctx || !p) return set_error(OVG_EINVAL, "NULL argument");
    *p = ctx->c.params;
    return OVG_OK;
}

// Validate topology: FREE slots have both links; PINNED / PINNED_SLIP slots
// (on the wall pin_wall, src/tangle.c:170-185) may have an open end; every
// link that exists is consistent.
static int validate_topology(int n, const int *conn, const int *status)
{
    for (int k = 0; k < n; ++k) {
        int s = status[2 * k];
        if (s == OVG_EMPTY) continue;
        const bool pinned = s == OVG_PINNED || s == OVG_PINNED_SLIP;
        if (!pinned && s != OVG_FREE) return set_error(OVG_EINVAL, "slot %d has status %d", k, s);
        if (pinned && (status[2 * k + 1] < 0 || status[2 * k + 1] > 5))
            return set_error(OVG_EINVAL, "pinned slot %d has pin_wall %d", k, status[2 * k + 1]);
        int f = conn[2 * k], r = conn[2 * k + 1];
        const bool f_open = pinned && f == -1, r_open = pinned && r == -1;
        if ((!f_open && (f < 0 || f >= n)) || (!r_open && (r < 0 || r >= n)))
            return set_error(OVG_EUNSUPPORTED,
                             "slot %d has an open end (forward %d, reverse %d) and is not "
                             "pinned: only closed loops and pinned lines are supported",
                             k, f, r);
        if ((!f_open && status[2 * f] == OVG_EMPTY) || (!r_open && status[2 * r] == OVG_EMPTY))
            return set_error(OVG_EINVAL, "slot %d links to an EMPTY slot", k);
        if ((!f_open && conn[2 * f + 1] != k) || (!r_open && conn[2 * r] != k))
            return set_error(OVG_EINVAL, "slot %d: forward/reverse links are inconsistent", k);
    }
    return OVG_OK;
}

static int alloc_state(Ctx &c, int n)
{
    size_t b8 = sizeof(double) * (size_t)(n > 0 ? n : 1);
    for (int q = 0; q < 3; ++q) {
        DBuf *bufs[] = {&c.x[q], &c.vels[q], &c.vs[q], &c.tan[q], &c.nor[q], &c.move[q],
                        &c.bs_prim[q], &c.bs_img[q], &c.xs[0][q], &c.xs[1][q]};
        for (auto *b : bufs) OVG_TRY(b->ensure(b8));
    }
    for (auto &s : c.seg) OVG_TRY(s.ensure(b8));
    OVG_TRY(c.seg_ids.ensure(sizeof(int2) * (size_t)(n > 0 ? n : 1)));
    size_t b4 = sizeof(int) * (size_t)(n > 0 ? n : 1);
    OVG_TRY(c.fwd.ensure(b4));
    OVG_TRY(c.rev.ensure(b4));
    OVG_TRY(c.status.ensure(b4));
    OVG_TRY(c.live.ensure(b4));
    OVG_TRY(c.src.ensure(b4));
    return OVG_OK;
}

static int upload_positions(Ctx &c, const double *vnodes)
{
    int n = c.n;
    c.rec_ok = false;
    OVG_TRY(positions_change(c));
    DBuf &stage = c.partial;
    OVG_TRY(stage.ensure(sizeof(double) * 3 * (size_t)n));
    OVG_HIP(hipMemcpyAsync(stage.p, vnodes, sizeof(double) * 3 * (size_t)n,
                           hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(n)), dim3(256), 0, c.stream, n,
                       stage.as<double>(), dp(c.x[0]), dp(c.x[1]), dp(c.x[2]));
    OVG_HIP(hipStreamSynchronize(c.stream));
    return OVG_OK;
}

static int upload_topology(Ctx &c, const int *conn, const int *status)
{
    int n = c.n;
    c.rec_ok = false;
    c.topo_version++;
    // the host's arrays replace whatever the device passes left
    c.dt_unread = false;
    c.live_pending = false;
    c.mirrors_stale = false;
    c.hint_on_device = false;
    if (conn != c.h_conn.data()) c.h_conn.assign(conn, conn + 2 * (size_t)n);
    if (status != c.h_status.data()) c.h_status.assign(status, status + 2 * (size_t)n);
    // the whole state is on the device now; every loop is new to the
    // small-loop pass
    c.dirty.clear();
    c.dirty_mark.assign((size_t)n, 0);
    c.new_pos.clear();
    c.new_xyz.clear();
    c.loop_dirty.clear();
    c.loop_dirty_mark.assign((size_t)n, 0);
    c.loops_full_scan = true;
    c.rc_where.assign((size_t)n, -1);
    c.rc_flag.assign((size_t)n, 0);
    std::vector<int> live;
    live.reserve(n);
    for (int k = 0; k < n; ++k)
        if (status[2 * k] != OVG_EMPTY) live.push_back(k);
    c.n_live = (int)live.size();
    c.n_src = c.n_live;  // closed loops: every live slot starts a segment
    DBuf &stage = c.partial;
    OVG_TRY(stage.ensure(sizeof(int) * 2 * (size_t)n));
    OVG_HIP(hipMemcpyAsync(stage.p, conn, sizeof(int) * 2 * (size_t)n, hipMemcpyHostToDevice,
                           c.stream));
    hipLaunchKernelGGL(k_split_pairs, dim3(grid_for(n)), dim3(256), 0, c.stream, n,
                       stage.as<int>(), c.fwd.as<int>(), c.rev.as<int>());
    OVG_HIP(hipStreamSynchronize(c.stream));
    std::vector<int> st(n);
    for (int k = 0; k < n; ++k) st[k] = status[2 * k];
    OVG_HIP(hipMemcpy(c.status.p, st.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
    // pin walls, kept only while the tangle has pinned nodes (the topology
    // passes, which would have to maintain them, refuse such tangles)
    c.has_pinned = false;
    for (int k = 0; k < n; ++k) {
        const bool pinned = st[k] == OVG_PINNED || st[k] == OVG_PINNED_SLIP;
        c.has_pinned |= pinned;
        st[k] = pinned ? status[2 * k + 1] : -1;
    }
    if (c.has_pinned) {
        OVG_TRY(c.pinw.ensure(sizeof(int) * (size_t)n));
        OVG_HIP(hipMemcpy(c.pinw.p, st.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
    }
    if (c.n_live) {
        OVG_HIP(hipMemcpy(c.live.p, live.data(), sizeof(int) * (size_t)c.n_live,
                          hipMemcpyHostToDevice));
        OVG_HIP(hipMemcpy(c.src.p, live.data(), sizeof(int) * (size_t)c.n_live,
                          hipMemcpyHostToDevice));
    }
    return OVG_OK;
}

extern "C" int ovg_upload(ovg_ctx *ctx, int n, const double *vnodes, const int *connections,
                          const int *status)
{
    clear_error();
    if (!ctx || n < 0 || (n > 0 && (!vnodes || !connections || !status)))
        return set_error(OVG_EINVAL, "bad arguments to ovg_upload");
    OVG_TRY(validate_topology(n, connections, status));
    for (int k = 0; k < n; ++k)
        if (status[2 * k] != OVG_EMPTY)
            for (int q = 0; q < 3; ++q)
                if (!std::isfinite(vnodes[3 * k + q]))
                    return set_error(OVG_EINVAL, "slot %d has a non-finite position", k);
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    OVG_TRY(alloc_state(c, n));
    if (c.n_tangles > 1) {  // a forest before: one tangle again
        c.n_tangles = 1;
        c.forest_params.clear();
        c.forest_off.clear();
        c.tid.release();
        push_params(c, 0.0);
    }
    c.n = n;
    c.next_empty = 0;
    OVG_TRY(upload_positions(c, vnodes));
    OVG_TRY(upload_topology(c, connections, status));
    for (int q = 0; q < 3; ++q) {
        DBuf *bufs[] = {&c.vels[q], &c.vs[q], &c.tan[q], &c.nor[q]};
        for (auto *b : bufs)
            hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n > 0 ? n : 1)), dim3(256), 0, c.stream,
                               n, b->as<double>());
    }
    OVG_HIP(hipStreamSynchronize(c.stream));
    return launch_status();
}

// The forest's parameter table: every tangle's constants with the context's
// launch settings (leaf size, group level: the autotuner changes them) and
// its fixed-point units; written again when those settings change.
int ovg::forest_table(Ctx &c, bool force)
{
    const int lf = c.params.leaf_size, gl = c.params.group_level;
    if (!force && lf == c.forest_leaf && gl == c.forest_glevel) return OVG_OK;
    const int K = c.n_tangles;
    std::vector<TangleParams> tab((size_t)K);
    for (int j = 0; j < K; ++j) {
        ovg_params q = c.forest_params[j];
        q.leaf_size = lf;
        q.group_level = gl;
        tab[j] = host_tparams(q, 0.0);
        double Lmax = 0;
        for (int a = 0; a < 3; ++a) Lmax = std::max(Lmax, q.box_hi[a] - q.box_lo[a]);
        fix_units(c.forest_live[j], Lmax, tab[j].fix_scale, tab[j].fix_inv, tab[j].fix_lim);
        tab[j].n_tangles = K;
        tab[j].tid = c.tid.as<int>();
    }
    OVG_TRY(c.tparams.ensure(sizeof(TangleParams) * (size_t)K));
    OVG_HIP(hipMemcpyAsync(c.tparams.p, tab.data(), sizeof(TangleParams) * (size_t)K,
                           hipMemcpyHostToDevice, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));  // tab is released on return
    c.forest_leaf = lf;
    c.forest_glevel = gl;
    return OVG_OK;
}

extern "C" int ovg_upload_forest(ovg_ctx *ctx, int n_tangles, const ovg_params *params,
                                 const int *slot_offsets, const double *vnodes,
                                 const int *connections, const int *status)
{
    clear_error();
    if (!ctx || n_tangles < 1 || !params || !slot_offsets)
        return set_error(OVG_EINVAL, "bad arguments to ovg_upload_forest");
    if (n_tangles > OVG_FOREST_MAX_TANGLES)
        return set_error(OVG_ELIMIT, "%d tangles exceed OVG_FOREST_MAX_TANGLES (%d)", n_tangles,
                         OVG_FOREST_MAX_TANGLES);
    if (slot_offsets[0] != 0) return set_error(OVG_EINVAL, "slot_offsets[0] must be 0");
    for (int j = 0; j < n_tangles; ++j)
        if (slot_offsets[j + 1] < slot_offsets[j])
            return set_error(OVG_EINVAL, "slot_offsets must not decrease (tangle %d)", j);
    const int n = slot_offsets[n_tangles];
    if (n > 0 && (!vnodes || !connections || !status))
        return set_error(OVG_EINVAL, "bad arrays for ovg_upload_forest");
    const ovg_params &p0 = params[0];
    for (int j = 0; j < n_tangles; ++j) {
        const ovg_params &q = params[j];
        const int r = validate_params(&q);
        if (r != OVG_OK) return set_error(r, "tangle %d: %s", j, ovg_last_error());
        if (q.method != OVG_METHOD_BH || q.traverse != OVG_TRAVERSE_LIST || q.reuse_lists)
            return set_error(OVG_EUNSUPPORTED, "tangle %d: a forest runs the Barnes-Hut list "
                             "traversal (method BH, traverse LIST, no reuse_lists)", j);
        if (q.precision != p0.precision || q.leaf_size != p0.leaf_size ||
            q.points_per_wave != p0.points_per_wave || q.group_level != p0.group_level ||
            q.rk4_classic != p0.rk4_classic || q.list_stack != p0.list_stack ||
            q.list_tile != p0.list_tile || q.list_blocks != p0.list_blocks)
            return set_error(OVG_EINVAL, "tangle %d: precision, leaf_size, points_per_wave, "
                             "group_level, rk4_classic and the list launch shape are one "
                             "launch's and must match tangle 0's", j);
        if (flow_depends_on_time(q))
            return set_error(OVG_EUNSUPPORTED, "tangle %d: a forest's parameter table is "
                             "written once, so its flows must not depend on time", j);
    }
    // per tangle: the reference layout with its own slot numbering
    std::vector<int> conn(2 * (size_t)n), st(2 * (size_t)n), tids((size_t)(n > 0 ? n : 1));
    std::vector<int> live_count((size_t)n_tangles, 0);
    for (int j = 0; j < n_tangles; ++j) {
        const int o = slot_offsets[j], m = slot_offsets[j + 1] - o;
        const int r = validate_topology(m, connections + 2 * (size_t)o, status + 2 * (size_t)o);
        if (r != OVG_OK) return set_error(r, "tangle %d: %s", j, ovg_last_error());
        for (int k = 0; k < m; ++k) {
            const size_t g = (size_t)o + k;
            for (int e = 0; e < 2; ++e) {
                const int l = connections[2 * g + e];
                conn[2 * g + e] = l < 0 ? -1 : l + o;
            }
            st[2 * g] = status[2 * g];
            st[2 * g + 1] = status[2 * g + 1];
            tids[g] = j;
            if (status[2 * g] != OVG_EMPTY) {
                ++live_count[j];
                for (int q = 0; q < 3; ++q)
                    if (!std::isfinite(vnodes[3 * g + q]))
                        return set_error(OVG_EINVAL, "tangle %d slot %d has a non-finite "
                                         "position", j, k);
            }
        }
    }
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    OVG_TRY(alloc_state(c, n));
    c.params = p0;
    c.have_params = true;
    c.rec_ok = false;
    c.n = n;
    c.next_empty = 0;
    c.n_tangles = n_tangles;
    c.forest_params.assign(params, params + n_tangles);
    c.forest_off.assign(slot_offsets, slot_offsets + n_tangles + 1);
    c.forest_live = live_count;
    OVG_TRY(c.tid.ensure(sizeof(int) * tids.size()));
    OVG_HIP(hipMemcpy(c.tid.p, tids.data(), sizeof(int) * tids.size(), hipMemcpyHostToDevice));
    OVG_TRY(forest_table(c, true));
    OVG_TRY(upload_positions(c, vnodes));
    OVG_TRY(upload_topology(c, conn.data(), st.data()));
    for (int q = 0; q < 3; ++q) {
        DBuf *bufs[] = {&c.vels[q], &c.vs[q], &c.tan[q], &c.nor[q]};
        for (auto *b : bufs)
            hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n > 0 ? n : 1)), dim3(256), 0, c.stream,
                               n, b->as<double>());
    }
    OVG_HIP(hipStreamSynchronize(c.stream));
    return launch_status();
}

extern "C" int ovg_forest_size(const ovg_ctx *ctx) { return ctx ? ctx->c.n_tangles : 0; }

// ---- HIP graphs of call sequences -------------------------------------------
extern "C" int ovg_graph_begin(ovg_ctx *ctx)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    Ctx &c = ctx->c;
    if (c.capturing) return set_error(OVG_EINVAL, "a capture is already open on this context");
    if (c.timing)
        return set_error(OVG_EINVAL, "timing events cannot be captured: ovg_set_timing(ctx, 0)");
    bool timed = flow_depends_on_time(c.params);
    for (const ovg_params &q : c.forest_params) timed |= flow_depends_on_time(q);
    if (timed)
        return set_error(OVG_EUNSUPPORTED, "a graph replays its parameters as captured, so the "
                         "flows must not depend on time");
    hipSetDevice(c.device);
    OVG_HIP(hipStreamSynchronize(c.stream));
    OVG_TRY(positions_change(c));  // the captured work moves points when it runs
    OVG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    c.capturing = true;
    t_capturing = true;
    t_capture_failed = false;
    return OVG_OK;
}

extern "C" int ovg_graph_end(ovg_ctx *ctx, int *n_nodes)
{
    clear_error();
    if (!ctx) return set_error(OVG_EINVAL, "context is NULL");
    Ctx &c = ctx->c;
    if (!c.capturing) return set_error(OVG_EINVAL, "no capture is open on this context");
    hipSetDevice(c.device);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c.stream, &g);
    c.capturing = false;
    t_capturing = false;
    (void)hipGetLastError();
    if (e != hipSuccess || !g)
        return set_error(OVG_EDEVICE, "the capture failed (%s): a call in it synchronised or "
                         "copied to the host", hipGetErrorString(e));
    if (t_capture_failed) {
        hipGraphDestroy(g);
        t_capture_failed = false;
        return set_error(OVG_EINVAL, "a call during the capture failed (a buffer had to grow): "
                         "nothing was instantiated");
    }
    if (c.graph_exec) {
        hipGraphExecDestroy(c.graph_exec);
        c.graph_exec = nullptr;
    }
    size_t nodes = 0;
    hipGraphGetNodes(g, nullptr, &nodes);
    const hipError_t ei = hipGraphInstantiate(&c.graph_exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) {
        c.graph_exec = nullptr;
        return set_error(OVG_EDEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
    }
    if (n_nodes) *n_nodes = (int)nodes;
    c.rec_ok = false;
    return launch_status();
}

extern "C" int ovg_graph_launch(ovg_ctx *ctx, int repeats)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    Ctx &c = ctx->c;
    if (!c.graph_exec) return set_error(OVG_EINVAL, "no graph has been captured on this context");
    if (repeats < 0) return set_error(OVG_EINVAL, "repeats must be >= 0");
    hipSetDevice(c.device);
    OVG_TRY(positions_change(c));
    for (int r = 0; r < repeats; ++r) OVG_HIP(hipGraphLaunch(c.graph_exec, c.stream));
    c.rec_ok = false;
    return launch_status();
}

extern "C" int ovg_graph_release(ovg_ctx *ctx)
{
    clear_error();
    if (!ctx) return set_error(OVG_EINVAL, "context is NULL");
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    if (c.graph_exec) {
        OVG_HIP(hipStreamSynchronize(c.stream));
        hipGraphExecDestroy(c.graph_exec);
        c.graph_exec = nullptr;
    }
    return OVG_OK;
}

extern "C" int ovg_upload_positions(ovg_ctx *ctx, const double *vnodes)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    return launch_status(upload_positions(c, vnodes));
}

extern "C" int ovg_upload_connections(ovg_ctx *ctx, const int *connections, const int *status)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_mirrors(ctx->c));
    OVG_TRY(single_tangle(ctx->c, __func__));
    Ctx &c = ctx->c;
    OVG_TRY(validate_topology(c.n, connections, status));
    hipSetDevice(c.device);
    return launch_status(upload_topology(c, connections, status));
}

extern "C" int ovg_upload_field(ovg_ctx *ctx, int field, const double *values)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    if (!values) return set_error(OVG_EINVAL, "values is NULL");
    Ctx &c = ctx->c;
    DBuf *dst;
    switch (field) {
    case OVG_FIELD_VELS: dst = c.vels; break;
    case OVG_FIELD_VS: dst = c.vs; break;
    case OVG_FIELD_TANGENTS: dst = c.tan; break;
    case OVG_FIELD_NORMALS: dst = c.nor; break;
    default: return set_error(OVG_EINVAL, "unknown field %d", field);
    }
    hipSetDevice(c.device);
    int n = c.n;
    OVG_TRY(c.partial.ensure(sizeof(double) * 3 * (size_t)n));
    OVG_HIP(hipMemcpyAsync(c.partial.p, values, sizeof(double) * 3 * (size_t)n,
                           hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(n)), dim3(256), 0, c.stream, n,
                       c.partial.as<double>(), dp(dst[0]), dp(dst[1]), dp(dst[2]));
    OVG_HIP(hipStreamSynchronize(c.stream));
    return launch_status();
}

static int check_flags(Ctx &c)
{
    int f = 0;
    OVG_HIP(hipMemcpyAsync(&f, c.flags.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    if (f) {
        OVG_HIP(hipMemsetAsync(c.flags.p, 0, sizeof(int), c.stream));
        if (f & 4)
            return set_error(OVG_EINVAL, "a target slot given to ovg_velocity_at_noalloc is out "
                                         "of range or EMPTY (its outputs are NaN)");
        if (f & 2) return set_error(OVG_EINVAL, "non-finite positions in the tangle state");
        return set_error(OVG_EINVAL, "a point left the box by more than 256 periods");
    }
    return OVG_OK;
}

extern "C" int ovg_download(ovg_ctx *ctx, double *vnodes, double *vels, double *vs,
                            double *tangents, double *normals)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    int n = c.n;
    OVG_TRY(c.partial.ensure(sizeof(double) * 3 * (size_t)n));
    double *outs[5] = {vnodes, vels, vs, tangents, normals};
    DBuf *src[5] = {c.x, c.vels, c.vs, c.tan, c.nor};
    for (int o = 0; o < 5; ++o) {
        if (!outs[o]) continue;
        hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(n)), dim3(256), 0, c.stream, n,
                           dp(src[o][0]), dp(src[o][1]), dp(src[o][2]), c.partial.as<double>());
        OVG_HIP(hipMemcpyAsync(outs[o], c.partial.p, sizeof(double) * 3 * (size_t)n,
                               hipMemcpyDeviceToHost, c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
    }
    return launch_status(check_flags(c));
}

extern "C" int ovg_download_connections(ovg_ctx *ctx, int *connections, int *status)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_mirrors(ctx->c));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    int n = c.n;
    if (connections) {
        OVG_TRY(c.partial.ensure(sizeof(int) * 2 * (size_t)n));
        hipLaunchKernelGGL(k_join_pairs, dim3(grid_for(n)), dim3(256), 0, c.stream, n,
                           c.fwd.as<int>(), c.rev.as<int>(), c.partial.as<int>());
        OVG_HIP(hipMemcpyAsync(connections, c.partial.p, sizeof(int) * 2 * (size_t)n,
                               hipMemcpyDeviceToHost, c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
    }
    if (status) std::memcpy(status, c.h_status.data(), sizeof(int) * 2 * (size_t)n);
    return launch_status();
}

extern "C" int ovg_num_slots(const ovg_ctx *ctx) { return ctx ? ctx->c.n : -1; }

// The live count after device topology passes is read from the device
// (waits); a failed read reports 0 with the error in ovg_last_error.
extern "C" int ovg_num_points(const ovg_ctx *ctx)
{
    if (!ctx) return 0;
    Ctx &c = const_cast<ovg_ctx *>(ctx)->c;
    if (c.dt_unread && sync_live(c) != OVG_OK) return 0;
    return c.n_live;
}

extern "C" int ovg_update_tangents_normals(ovg_ctx *ctx)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    Triple x = tri(c.x), t = tri(c.tan), nrm = tri(c.nor);
    launch_stencil(c, x.v, t.v, nrm.v);
    return launch_status();
}

static int velocities(Ctx &c, double t, bool own_tangents)
{
    push_params(c, t);
    Triple x = tri(c.x);
    OVG_TRY(eval_bs(c, x.v, own_tangents ? kEvalStep : kEvalPlain));
    launch_compose(c, x.v, own_tangents, 0, 0.0);
    return OVG_OK;
}

extern "C" int ovg_update_velocities(ovg_ctx *ctx, double t)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    return launch_status(velocities(c, t, false));
}

extern "C" int ovg_update_tangle(ovg_ctx *ctx, double t)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    return launch_status(velocities(c, t, true));
}

extern "C" int ovg_rk4_step(ovg_ctx *ctx, double t, double dt)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    double h4 = c.params.rk4_classic ? dt : dt / 2;
    (void)h4;
    OVG_TRY(positions_change(c));  // the step moves every point
    // stage positions: xs[0] = x + dt/2 k1 (euler_step2, src/vortex_dynamics.c:76)
    Triple k1 = tri(c.vels), xs0 = tri(c.xs[0]), xs1 = tri(c.xs[1]);
    launch_stage_positions(c, k1.v, dt / 2, xs0.v);
    double times[3] = {t + dt / 2, t + dt / 2, t + dt};
    for (int s = 1; s <= 3; ++s) {
        const double *const *xe = (s == 2) ? xs1.v : xs0.v;
        push_params(c, times[s - 1]);
        const int r = eval_bs(c, xe, kEvalStage);
        if (r != OVG_OK) {
            c.rec_ok = false;
            return r;
        }
        launch_compose(c, xe, true, s, dt);
    }
    c.rec_ok = false;  // recorded lists serve one step
    return launch_status();
}

extern "C" int ovg_enforce_periodic(ovg_ctx *ctx)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    launch_wrap(c);
    return launch_status();
}

extern "C" int ovg_velocity_at(ovg_ctx *ctx, int m, const int *slots, double t, double *vs,
                               double *vels)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_mirrors(ctx->c));
    OVG_TRY(single_tangle(ctx->c, __func__));
    Ctx &c = ctx->c;
    if (m < 0 || (m > 0 && !slots)) return set_error(OVG_EINVAL, "bad slot list");
    for (int i = 0; i < m; ++i)
        if (slots[i] < 0 || slots[i] >= c.n || c.h_status[2 * slots[i]] == OVG_EMPTY)
            return set_error(OVG_EINVAL, "slot %d is not a live point", slots[i]);
    if (m == 0) return OVG_OK;
    hipSetDevice(c.device);
    push_params(c, t);
    Triple x = tri(c.x);
    launch_prep_segments(c, x.v);
    OVG_TRY(c.rc_tmp.ensure(sizeof(int) * (size_t)m));
    OVG_HIP(hipMemcpyAsync(c.rc_tmp.p, slots, sizeof(int) * (size_t)m, hipMemcpyHostToDevice,
                           c.stream));
    Triple bp = tri(c.bs_prim), bi = tri(c.bs_img);
    ovg_params saved = c.params;
    c.params.precision = OVG_PREC_FP64;
    const int rd = launch_direct(c, x.v, m, c.rc_tmp.as<int>(), bp.v, bi.v, 0);
    c.params = saved;
    OVG_TRY(rd);
    // compose over the selected slots only
    int saved_live = c.n_live;
    DBuf saved_list = c.live;
    c.live = c.rc_tmp;
    c.n_live = m;
    launch_compose(c, x.v, false, 0, 0.0);
    c.live = saved_list;
    c.n_live = saved_live;
    std::vector<double> hv(3 * (size_t)c.n);
    for (int o = 0; o < 2; ++o) {
        double *out = o == 0 ? vs : vels;
        if (!out) continue;
        DBuf *src = o == 0 ? c.vs : c.vels;
        OVG_TRY(c.partial.ensure(sizeof(double) * 3 * (size_t)c.n));
        hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(c.n)), dim3(256), 0, c.stream, c.n,
                           dp(src[0]), dp(src[1]), dp(src[2]), c.partial.as<double>());
        OVG_HIP(hipMemcpyAsync(hv.data(), c.partial.p, sizeof(double) * 3 * (size_t)c.n,
                               hipMemcpyDeviceToHost, c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
        for (int i = 0; i < m; ++i)
            for (int q = 0; q < 3; ++q) out[3 * i + q] = hv[3 * (size_t)slots[i] + q];
    }
    return launch_status();
}

// out[3 i + q] = v_q[slot_i] (AoS), NaN for a slot of -1
__global__ void k_gather_aos(int m, const int *__restrict__ slots, const double *__restrict__ vx,
                             const double *__restrict__ vy, const double *__restrict__ vz,
                             double *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int k = slots[i];
    const double nan = __builtin_nan("");
    out[3 * i] = k < 0 ? nan : vx[k];
    out[3 * i + 1] = k < 0 ? nan : vy[k];
    out[3 * i + 2] = k < 0 ? nan : vz[k];
}

extern "C" int ovg_velocity_at_noalloc(ovg_ctx *ctx, int m, const int *d_slots, double t,
                                       double *d_vs, double *d_vels)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    OVG_TRY(single_tangle(ctx->c, __func__));
    Ctx &c = ctx->c;
    if (m < 0 || (m > 0 && !d_slots)) return set_error(OVG_EINVAL, "bad slot list");
    if (m > OVG_NOALLOC_MAX_TARGETS)
        return set_error(OVG_ELIMIT,
                         "ovg_velocity_at_noalloc: %d targets exceed its maximum of %d "
                         "(OVG_NOALLOC_MAX_TARGETS); ovg_velocity_at takes any number",
                         m, OVG_NOALLOC_MAX_TARGETS);
    if (m == 0) return OVG_OK;
    hipSetDevice(c.device);
    push_params(c, t);
    Triple x = tri(c.x);
    launch_prep_segments(c, x.v);
    Triple bp = tri(c.bs_prim), bi = tri(c.bs_img);
    ovg_params saved = c.params;
    c.params.precision = OVG_PREC_FP64;
    const int *checked = nullptr;
    const int rd = launch_direct_noalloc(c, x.v, m, d_slots, bp.v, bi.v, &checked);
    c.params = saved;
    OVG_TRY(rd);
    // compose over the checked slots only (a view of the scratch list: the
    // context's own live list is restored, nothing is allocated or freed)
    const int saved_live = c.n_live;
    const DBuf saved_list = c.live;
    c.live.p = const_cast<int *>(checked);
    c.live.bytes = sizeof(int) * (size_t)m;
    c.n_live = m;
    launch_compose(c, x.v, false, 0, 0.0);
    c.live = saved_list;
    c.n_live = saved_live;
    const int gs = grid_for(m);
    if (d_vs)
        hipLaunchKernelGGL(k_gather_aos, dim3(gs), dim3(256), 0, c.stream, m, checked,
                           dp(c.vs[0]), dp(c.vs[1]), dp(c.vs[2]), d_vs);
    if (d_vels)
        hipLaunchKernelGGL(k_gather_aos, dim3(gs), dim3(256), 0, c.stream, m, checked,
                           dp(c.vels[0]), dp(c.vels[1]), dp(c.vels[2]), d_vels);
    return launch_status();
}

extern "C" int ovg_line_length(ovg_ctx *ctx, double *length, int *n_points)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    int gs = grid_for(c.n_live);
    std::vector<double> part(gs > 0 ? gs : 1, 0.0);
    if (gs > 0) {
        OVG_TRY(c.partial.ensure(sizeof(double) * (size_t)gs));
        launch_line_length(c, c.partial.as<double>());
        OVG_HIP(hipMemcpyAsync(part.data(), c.partial.p, sizeof(double) * (size_t)gs,
                               hipMemcpyDeviceToHost, c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
    }
    double L = 0;
    for (int b = 0; b < gs; ++b) L += part[b];
    if (length) *length = L;
    if (n_points) *n_points = c.n_live;
    return launch_status();
}

extern "C" int ovg_get_stats(ovg_ctx *ctx, ovg_stats *s)
{
    clear_error();
    if (!ctx || !s) return set_error(OVG_EINVAL, "NULL argument");
    Ctx &c = ctx->c;
    OVG_TRY(sync_live(c));
    hipSetDevice(c.device);
    return launch_status(fetch_stats(c, s));
}

extern "C" int ovg_set_timing(ovg_ctx *ctx, int on)
{
    if (!ctx) return set_error(OVG_EINVAL, "context is NULL");
    ctx->c.timing = on;
    return launch_status();
}

extern "C" int ovg_synchronize(ovg_ctx *ctx)
{
    clear_error();
    if (!ctx) return set_error(OVG_EINVAL, "context is NULL");
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    OVG_HIP(hipStreamSynchronize(c.stream));
    return check_flags(c);
}

extern "C" int ovg_marker_record(ovg_ctx *ctx, int id)
{
    clear_error();
    if (!ctx) return set_error(OVG_EINVAL, "context is NULL");
    if (id < 0 || id >= OVG_MAX_MARKERS)
        return set_error(OVG_EINVAL, "marker id %d outside [0, %d)", id, OVG_MAX_MARKERS);
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    if ((int)c.markers.size() <= id) c.markers.resize((size_t)id + 1, nullptr);
    if (!c.markers[id]) OVG_HIP(hipEventCreate(&c.markers[id]));
    OVG_HIP(hipEventRecord(c.markers[id], c.stream));
    return launch_status();
}

extern "C" int ovg_marker_elapsed(ovg_ctx *ctx, int id0, int id1, double *ms)
{
    clear_error();
    if (!ctx || !ms) return set_error(OVG_EINVAL, "context or output is NULL");
    Ctx &c = ctx->c;
    for (int id : {id0, id1})
        if (id < 0 || id >= (int)c.markers.size() || !c.markers[id])
            return set_error(OVG_EINVAL, "marker %d was never recorded", id);
    hipSetDevice(c.device);
    OVG_HIP(hipEventSynchronize(c.markers[id1]));
    float f = 0.f;
    OVG_HIP(hipEventElapsedTime(&f, c.markers[id0], c.markers[id1]));
    *ms = f;
    return launch_status();
}

extern "C" int ovg_reconnect_candidates(ovg_ctx *ctx, int max_pairs, int *pairs, int *n_pairs)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    OVG_TRY(single_tangle(ctx->c, __func__));
    if (max_pairs < 0 || (max_pairs > 0 && !pairs))
        return set_error(OVG_EINVAL, "bad pair buffer");
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    return launch_status(recon_candidates(c, max_pairs, pairs, n_pairs));
}

extern "C" int ovg_reconnect(ovg_ctx *ctx, double t, int *n_reconnections)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(single_tangle(ctx->c, __func__));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    if (devtopo_applies(c)) {
        OVG_TRY(devtopo_reconnect(c, t));
        if (n_reconnections) {
            int cnt[4];
            OVG_TRY(devtopo_counts(c, cnt));
            *n_reconnections = cnt[0];
        }
        return launch_status();
    }
    OVG_TRY(sync_mirrors(c));
    int nrec = 0;
    OVG_TRY(reconnect_pass(c, t, &nrec));
    if (n_reconnections) *n_reconnections = nrec;
    return launch_status();
}

extern "C" int ovg_topology_counters(ovg_ctx *ctx, int *out)
{
    clear_error();
    OVG_TRY(check_ready(ctx, false));
    if (!out) return set_error(OVG_EINVAL, "out is NULL");
    hipSetDevice(ctx->c.device);
    return launch_status(devtopo_counters(ctx->c, out));
}

extern "C" int ovg_smooth(ovg_ctx *ctx)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    OVG_TRY(single_tangle(ctx->c, __func__));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    return launch_status(smooth_loops(c));
}

// eliminate_small_loops, src/tangle.c:1063-1128 (closed loops).  Loops are
// disjoint and each is deleted exactly when it has fewer than cutoff
// points, so the result does not depend on the visiting order.  A loop can
// only fall below the cutoff through a topology change, so after the first
// full pass (and while the cutoff does not grow) only the loops through
// slots touched since the previous pass are walked, each for at most
// cutoff steps.
extern "C" int ovg_eliminate_small_loops(ovg_ctx *ctx, int cutoff, int *n_removed)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(single_tangle(ctx->c, __func__));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    if (devtopo_applies(c)) {
        OVG_TRY(devtopo_small_loops(c, cutoff));
        if (n_removed) {
            int cnt[4];
            OVG_TRY(devtopo_counts(c, cnt));
            *n_removed = cnt[1];
        }
        return launch_status();
    }
    OVG_TRY(sync_mirrors(c));
    int n = c.n;
    std::vector<int> &conn = c.h_conn, &st = c.h_status;
    int removed = 0;
    auto erase_loop = [&](int start) {
        int next = start;
        for (;;) {
            int tmp = next;
            next = conn[2 * (size_t)next];
            conn[2 * (size_t)tmp] = -1;
            conn[2 * (size_t)tmp + 1] = -1;
            st[2 * (size_t)tmp] = OVG_EMPTY;
            mark_dirty(c, tmp);
            removed++;
            if (next == start || next < 0) break;
        }
    };
    if (c.has_pinned) {
        // open lines ending on pinned nodes: the reference's full pass
        // (src/tangle.c:1063-1128).  The forward walk that reaches a pinned
        // end restarts at k and adds the steps back to the other end; the
        // erase then runs forward from that end.  Every pass is a full one
        // while pinned lines exist.
        std::vector<char> visited(n, 0);
        for (int k = 0; k < n; ++k) {
            if (st[2 * (size_t)k] == OVG_EMPTY || visited[k]) continue;
            visited[k] = 1;
            int loop = 0, here = k, next = conn[2 * (size_t)k];
            while (next != k) {
                const int sh = st[2 * (size_t)here];
                if ((sh == OVG_PINNED || sh == OVG_PINNED_SLIP) && next < 0) {
                    here = k;
                    next = conn[2 * (size_t)here + 1];
                    while (next >= 0) {
                        visited[here] = 1;
                        here = next;
                        next = conn[2 * (size_t)here + 1];
                        loop++;
                    }
                    break;
                }
                visited[here] = 1;
                here = next;
                next = conn[2 * (size_t)here];
                loop++;
            }
            if (loop < cutoff) erase_loop(here);
        }
        c.loops_full_scan = true;
        // when the pass erased every pinned line the tangle is closed loops
        // again: the topology passes take it and pin walls are dropped
        bool pinned_left = false;
        for (int k = 0; k < n && !pinned_left; ++k) {
            const int s = st[2 * (size_t)k];
            pinned_left = s == OVG_PINNED || s == OVG_PINNED_SLIP;
        }
        if (!pinned_left) {
            c.has_pinned = false;
            c.pinw.release();
        }
    } else if (c.loops_full_scan || cutoff > c.small_cutoff_seen) {
        std::vector<char> visited(n, 0);
        for (int k = 0; k < n; ++k) {
            if (st[2 * (size_t)k] == OVG_EMPTY || visited[k]) continue;
            visited[k] = 1;
            int loop = 0, here = k, next = conn[2 * (size_t)k];
            while (next != k) {
                visited[here] = 1;
                here = next;
                next = conn[2 * (size_t)here];
                loop++;
            }
            if (loop < cutoff) erase_loop(here);
        }
        c.loops_full_scan = false;
        c.small_cutoff_seen = cutoff;
    } else {
        for (size_t i = 0; i < c.loop_dirty.size(); ++i) {
            const int k = c.loop_dirty[i];
            if (k >= n || st[2 * (size_t)k] == OVG_EMPTY) continue;
            int loop = 0, here = k, next = conn[2 * (size_t)k];
            while (next != k && loop < cutoff) {
                here = next;
                next = conn[2 * (size_t)here];
                loop++;
            }
            if (next == k && loop < cutoff) erase_loop(k);
        }
    }
    for (int k : c.loop_dirty)
        if ((size_t)k < c.loop_dirty_mark.size()) c.loop_dirty_mark[k] = 0;
    c.loop_dirty.clear();
    if (n_removed) *n_removed = removed;
    if (removed) {
        hipSetDevice(c.device);
        OVG_TRY(commit_topology(c));
        // the erased slots were marked for the next pass too; they are EMPTY
        for (int k : c.loop_dirty)
            if ((size_t)k < c.loop_dirty_mark.size()) c.loop_dirty_mark[k] = 0;
        c.loop_dirty.clear();
    }
    return launch_status();
}

// The host doubled the slot count (expand_tangle, src/tangle.c:68-101):
// reallocate, keep every old slot's vels, vs, tangent and normal (the
// reference's realloc keeps them; the new slots read zero here), and upload
// the positions and the topology.
static int grow_slots(Ctx &c, int n_new, const double *X)
{
    const int n_old = c.n;
    DBuf *f[12] = {&c.vels[0], &c.vels[1], &c.vels[2], &c.vs[0], &c.vs[1], &c.vs[2],
                   &c.tan[0],  &c.tan[1],  &c.tan[2],  &c.nor[0], &c.nor[1], &c.nor[2]};
    const size_t b_old = sizeof(double) * (size_t)n_old;
    DBuf keep;
    int r = keep.ensure(12 * b_old);
    for (int i = 0; i < 12 && r == OVG_OK; ++i)
        if (hipMemcpyAsync(keep.as<unsigned char>() + i * b_old, f[i]->p, b_old,
                           hipMemcpyDeviceToDevice, c.stream) != hipSuccess)
            r = set_error(OVG_EDEVICE, "state copy failed");
    if (r == OVG_OK) r = alloc_state(c, n_new);
    if (r == OVG_OK) {
        c.n = n_new;
        const size_t b_new = sizeof(double) * (size_t)n_new;
        for (int i = 0; i < 12 && r == OVG_OK; ++i)
            if (hipMemcpyAsync(f[i]->p, keep.as<unsigned char>() + i * b_old, b_old,
                               hipMemcpyDeviceToDevice, c.stream) != hipSuccess ||
                hipMemsetAsync(f[i]->as<unsigned char>() + b_old, 0, b_new - b_old,
                               c.stream) != hipSuccess)
                r = set_error(OVG_EDEVICE, "state copy failed");
    }
    if (r == OVG_OK) r = upload_positions(c, X);  // synchronises the copies above
    if (r == OVG_OK) r = upload_topology(c, c.h_conn.data(), c.h_status.data());
    hipStreamSynchronize(c.stream);
    keep.release();
    return r;
}

int ovg::grow_capacity(Ctx &c, int n_new)
{
    const int n_old = c.n;
    if (n_new <= n_old) return OVG_OK;
    const double *Xp = nullptr;
    OVG_TRY(host_positions(c, &Xp));
    std::vector<double> X(3 * (size_t)n_new, 0.0);
    std::memcpy(X.data(), Xp, sizeof(double) * 3 * (size_t)n_old);
    c.h_conn.resize(2 * (size_t)n_new, -1);
    c.h_status.resize(2 * (size_t)n_new, -1);
    return grow_slots(c, n_new, X.data());
}

// ---- remesh (src/tangle.c:1018-1061) on the host mirrors ----------------
namespace {

struct HostTangle {
    int N;
    double *X;                   // AoS positions: the pinned mirror, or Xv
    std::vector<double> Xv;      // owns the positions once the tangle grew
    std::vector<int> &conn, &st; // the context's mirrors, edited in place
    int next_empty;
    bool grown = false;

    HostTangle(int n, const double *x, std::vector<int> &cn, std::vector<int> &s, int ne)
        : N(n), X(const_cast<double *>(x)), conn(cn), st(s), next_empty(ne)
    {
    }

    // search_next_empty_node, src/tangle.c:201-234 (a read past the last
    // slot counts as not EMPTY)
    int search()
    {
        if (next_empty > 0) {
            if (next_empty < N && st[2 * (size_t)next_empty] == OVG_EMPTY) {
                int r = next_empty;
                if (next_empty + 1 < N && st[2 * (size_t)(next_empty + 1)] == OVG_EMPTY)
                    ++next_empty;
                else
                    next_empty = -1;
                return r;
            }
        }
        for (int k = 0; k < N; ++k)
            if (st[2 * (size_t)k] == OVG_EMPTY) {
                next_empty = (k + 1 < N && st[2 * (size_t)(k + 1)] == OVG_EMPTY) ? k + 1 : -1;
                return k;
            }
        return -1;
    }
    // expand_tangle(tangle, 2 N), src/tangle.c:68-101: the new slots are
    // EMPTY with no links, then next_empty = get_next_empty_node
    void expand()
    {
        int old = N;
        N = 2 * old;
        if (!grown) {
            Xv.assign(X, X + 3 * (size_t)old);
            grown = true;
        }
        Xv.resize(3 * (size_t)N, 0.0);
        X = Xv.data();
        conn.resize(2 * (size_t)N, -1);
        st.resize(2 * (size_t)N, -1);
        next_empty = search();
    }
    // get_next_empty_node :241-253
    int next_empty_node()
    {
        int idx = search();
        while (idx < 0 || st[2 * (size_t)idx] != OVG_EMPTY) {
            expand();
            idx = search();
        }
        return idx;
    }
    // number_of_empty_nodes :128-133
    int empty_count() const
    {
        int e = 0;
        for (int k = 0; k < N; ++k) e += st[2 * (size_t)k] == OVG_EMPTY;
        return e;
    }
};

// add_point, src/tangle.c:261-335 (st: status mirror when the tangle has
// pinned nodes, whose stencils step_node takes through the wall)
int add_point(HostTangle &h, const ovg::TangleParams &P, int p, const int *st)
{
    using namespace ovg::host;
    int nw = h.next_empty_node();
    h.st[2 * (size_t)nw] = OVG_FREE;
    h.st[2 * (size_t)nw + 1] = -1;
    int next = h.conn[2 * (size_t)p];
    double t0[3], n0[3], t1[3], n1[3];
    tangent_normal(P, h.X, h.conn.data(), p, t0, n0, st);
    tangent_normal(P, h.X, h.conn.data(), next, t1, n1, st);
    double s0[3] = {h.X[3 * (size_t)p], h.X[3 * (size_t)p + 1], h.X[3 * (size_t)p + 2]};
    double s1[3] = {h.X[3 * (size_t)next], h.X[3 * (size_t)next + 1], h.X[3 * (size_t)next + 2]};
    pwrap(P, s0, s1);
    double l = dist3(s0, s1);
    double nv[3], nwp[3];
    for (int c = 0; c < 3; ++c) nv[c] = n0[c] + n1[c];
    for (int c = 0; c < 3; ++c) nv[c] = nv[c] * 0.5;
    if (len3(nv) > 1e-5) {
        double R = 1 / len3(nv);
        double dR = R * R - l * l / 4;
        double delta = dR > 0 ? R - std::sqrt(dR) : R;
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
    for (int c = 0; c < 3; ++c) h.X[3 * (size_t)nw + c] = nwp[c];
    h.conn[2 * (size_t)nw + 1] = p;
    h.conn[2 * (size_t)nw] = next;
    h.conn[2 * (size_t)p] = nw;
    h.conn[2 * (size_t)next + 1] = nw;
    return nw;
}

}  // namespace

// tangents, normals, vs and vels of m slots from AoS records of 12 doubles
__global__ void k_scatter_state(int m, const int *__restrict__ slots,
                                const double *__restrict__ v, double *tx, double *ty, double *tz,
                                double *nx, double *ny, double *nz, double *sx, double *sy,
                                double *sz, double *ex, double *ey, double *ez)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int k = slots[i];
    const double *r = v + 12 * (size_t)i;
    tx[k] = r[0]; ty[k] = r[1]; tz[k] = r[2];
    nx[k] = r[3]; ny[k] = r[4]; nz[k] = r[5];
    sx[k] = r[6]; sy[k] = r[7]; sz[k] = r[8];
    ex[k] = r[9]; ey[k] = r[10]; ez[k] = r[11];
}

// Pin walls on the device from the host status mirror, kept while the
// tangle has pinned nodes.
static int refresh_pins(Ctx &c)
{
    std::vector<int> pw((size_t)c.n);
    bool any = false;
    for (int k = 0; k < c.n; ++k) {
        const int s = c.h_status[2 * (size_t)k];
        const bool pinned = s == OVG_PINNED || s == OVG_PINNED_SLIP;
        pw[k] = pinned ? c.h_status[2 * (size_t)k + 1] : -1;
        any |= pinned;
    }
    c.has_pinned = any;
    if (!any) {
        c.pinw.release();
        return OVG_OK;
    }
    c.loops_full_scan = true;
    OVG_TRY(c.pinw.ensure(sizeof(int) * (size_t)c.n));
    OVG_HIP(hipMemcpyAsync(c.pinw.p, pw.data(), sizeof(int) * (size_t)c.n, hipMemcpyHostToDevice,
                           c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    return OVG_OK;
}

// enforce_wall_boundaries, src/tangle.c:966-1016, on the host mirrors.
// Wall by wall (solid ones: not OPEN, not PERIODIC) and in slot order, every
// FREE point within rec_dist of the wall (check_wall, :822-840) that no clip
// of this pass has flagged gets its tangent and its direct-sum velocity
// (update_tangent_normal, update_velocity with tree == NULL) and is clipped
// to the wall (connect_to_wall, :876-958) unless it moves away from the wall,
// runs parallel to it, or the two new wall segments would be longer; then
// every FREE point behind a solid wall is removed (:996-1015).
// A clip moves no existing point: the velocities are the direct sums at the
// pass's start, computed for every candidate in one launch, plus the exact
// change (all images, mirrors unreflected as the direct path has them) of
// each segment the clips so far replaced or added.
static int walls_pass(Ctx &c, double t, int *n_clip, int *n_rem)
{
    *n_clip = *n_rem = 0;
    const ovg_params &pr = c.params;
    bool solid[6], any = false;
    for (int f = 0; f < 6; ++f) {
        solid[f] = pr.wall[f] != OVG_WALL_OPEN && pr.wall[f] != OVG_WALL_PERIODIC;
        any |= solid[f];
    }
    if (!any || c.n_live == 0) return OVG_OK;
    const double *Xp = nullptr;
    OVG_TRY(host_positions(c, &Xp));
    HostTangle h(c.n, Xp, c.h_conn, c.h_status, c.next_empty);
    const TangleParams P = host_tparams(pr, t);
    const double rd = pr.rec_dist;
    auto free_at = [&](int k) { return h.st[2 * (size_t)k] == OVG_FREE; };
    // wall_dist, :848-866 (axis of face w: w / 2; LEFT, BACK, DOWN even)
    auto wall_dist = [&](int k, int w) {
        const int a = w >> 1;
        const double x = h.X[3 * (size_t)k + a];
        return (w & 1) ? pr.box_hi[a] - x : x - pr.box_lo[a];
    };
    // check_wall, :822-840
    auto check_wall = [&](int k, int w) {
        const int a = w >> 1;
        const double x = h.X[3 * (size_t)k + a];
        return (w & 1) ? x > pr.box_hi[a] - rd : x < rd + pr.box_lo[a];
    };
    std::vector<int> cand;
    for (int k = 0; k < c.n; ++k) {
        if (!free_at(k)) continue;
        bool near = false;
        for (int w = 0; w < 6 && !near; ++w) near = solid[w] && check_wall(k, w);
        if (near) cand.push_back(k);
    }
    std::vector<double> bsp, bsi;
    std::vector<int> where((size_t)c.n, -1);
    const int m = (int)cand.size();
    if (m) {
        for (int i = 0; i < m; ++i) where[cand[i]] = i;
        push_params(c, t);
        Triple x = tri(c.x);
        launch_prep_segments(c, x.v);
        OVG_TRY(c.rc_tmp.ensure(sizeof(int) * (size_t)m));
        OVG_HIP(hipMemcpyAsync(c.rc_tmp.p, cand.data(), sizeof(int) * (size_t)m,
                               hipMemcpyHostToDevice, c.stream));
        Triple bp = tri(c.bs_prim), bi = tri(c.bs_img);
        const ovg_params saved = c.params;
        c.params.precision = OVG_PREC_FP64;
        const int r = launch_direct(c, x.v, m, c.rc_tmp.as<int>(), bp.v, bi.v, 0);
        c.params = saved;
        OVG_TRY(r);
        OVG_TRY(c.partial.ensure(sizeof(double) * 3 * (size_t)m));
        bsp.resize(3 * (size_t)m);
        bsi.resize(3 * (size_t)m);
        for (int o = 0; o < 2; ++o) {
            const Triple &src = o == 0 ? bp : bi;
            hipLaunchKernelGGL(k_gather_aos, dim3((m + 255) / 256), dim3(256), 0, c.stream, m,
                               c.rc_tmp.as<int>(), src.v[0], src.v[1], src.v[2],
                               c.partial.as<double>());
            OVG_HIP(hipMemcpyAsync((o == 0 ? bsp : bsi).data(), c.partial.p,
                                   sizeof(double) * 3 * (size_t)m, hipMemcpyDeviceToHost,
                                   c.stream));
            OVG_HIP(hipStreamSynchronize(c.stream));
        }
    }
    struct Change {
        int seg, old_f, new_f;
    };
    std::vector<Change> changes;
    std::vector<int> recalc((size_t)c.n, 0);
    std::vector<int> visited, vis_at((size_t)c.n, -1);  // slots update_velocity wrote
    std::vector<double> state;          // their tangent, normal, vs, vels
    std::vector<int> touched, inserted;
    const int pin = pr.pin_mode ? pr.pin_mode : OVG_PINNED;
    const double cut = std::cos(pr.rec_angle);
    auto connect = [&](int k, int w) -> bool {
        double tk[3], nk[3], vs[3], vels[3], bs[3];
        host::tangent_normal(P, h.X, h.conn.data(), k, tk, nk, h.st.data());
        const int e = where[k];
        for (int a = 0; a < 3; ++a) bs[a] = bsp[3 * (size_t)e + a] + bsi[3 * (size_t)e + a];
        for (const Change &ch : changes) {
            double vo[3] = {0, 0, 0}, vn[3] = {0, 0, 0};
            if (ch.old_f >= 0) host::segment_all_images(P, h.X, ch.seg, ch.old_f, k, vo, false);
            if (ch.new_f >= 0) host::segment_all_images(P, h.X, ch.seg, ch.new_f, k, vn, false);
            for (int a = 0; a < 3; ++a) bs[a] += vn[a] - vo[a];
        }
        host::compose(P, h.X, h.conn.data(), k, tk, nk, bs, vs, vels, h.st.data());
        // the last visit of a point (a corner point may be tested for two
        // walls) leaves its state
        if (vis_at[k] < 0) {
            vis_at[k] = (int)visited.size();
            visited.push_back(k);
            state.resize(state.size() + 12);
        }
        double *rec = &state[12 * (size_t)vis_at[k]];
        for (const double *v : {tk, nk, vs, vels})
            for (int a = 0; a < 3; ++a) *rec++ = v[a];
        double nw[3];
        host::wall_normal(w, nw);
        const double d0 = wall_dist(k, w);
        if (d0 > 0 && host::dot3(vels, nw) > 0) return false;  // moving away
        // parallel to the wall: the reference takes abs() of the cosine, an
        // int, so only |cos| >= 1 counts (:892)
        const double nd = host::dot3(tk, nw) / host::len3(tk) / host::len3(nw);
        if (std::abs((int)nd) > cut) return false;
        const int next = h.conn[2 * (size_t)k], prev = h.conn[2 * (size_t)k + 1];
        const double d1 = wall_dist(next, w), dm1 = wall_dist(prev, w);
        auto seg_len = [&](int a, int b) {
            double q[3] = {h.X[3 * (size_t)b], h.X[3 * (size_t)b + 1], h.X[3 * (size_t)b + 2]};
            host::pwrap(P, &h.X[3 * (size_t)a], q);
            return host::dist3(&h.X[3 * (size_t)a], q);
        };
        const double vd1 = seg_len(k, next), vdm1 = seg_len(k, prev);
        if (vd1 < d0 + d1 && vdm1 < d0 + dm1) return false;  // total length would grow
        const int p1 = h.next_empty_node();
        h.st[2 * (size_t)p1] = pin;
        h.st[2 * (size_t)p1 + 1] = w;
        const int p2 = h.next_empty_node();
        h.st[2 * (size_t)p2] = pin;
        h.st[2 * (size_t)p2 + 1] = w;
        if ((int)recalc.size() < h.N) recalc.resize((size_t)h.N, 0);
        for (int a = 0; a < 3; ++a)
            h.X[3 * (size_t)p1 + a] = h.X[3 * (size_t)k + a] + nw[a] * -d0;
        for (int sl : {k, p1, p2, next, prev}) recalc[sl]++;
        if (d1 < dm1) {
            for (int a = 0; a < 3; ++a)
                h.X[3 * (size_t)p2 + a] = h.X[3 * (size_t)next + a] + nw[a] * -d1;
            h.conn[2 * (size_t)k] = p1;
            h.conn[2 * (size_t)p1] = -1;
            h.conn[2 * (size_t)p1 + 1] = k;
            h.conn[2 * (size_t)p2] = next;
            h.conn[2 * (size_t)p2 + 1] = -1;
            h.conn[2 * (size_t)next + 1] = p2;
            changes.push_back({k, next, p1});
            changes.push_back({p2, -1, next});
        } else {
            for (int a = 0; a < 3; ++a)
                h.X[3 * (size_t)p2 + a] = h.X[3 * (size_t)prev + a] + nw[a] * -dm1;
            h.conn[2 * (size_t)k + 1] = p1;
            h.conn[2 * (size_t)p1] = k;
            h.conn[2 * (size_t)p1 + 1] = -1;
            h.conn[2 * (size_t)p2] = -1;
            h.conn[2 * (size_t)p2 + 1] = prev;
            h.conn[2 * (size_t)prev] = p2;
            changes.push_back({prev, k, p2});
            changes.push_back({p1, -1, k});
        }
        for (int sl : {k, next, prev, p1, p2}) touched.push_back(sl);
        inserted.push_back(p1);
        inserted.push_back(p2);
        return true;
    };
    int clips = 0, removed = 0;
    for (int w = 0; w < 6; ++w) {
        if (!solid[w]) continue;
        for (int k : cand) {
            if (!free_at(k) || recalc[k]) continue;
            if (check_wall(k, w) && connect(k, w)) ++clips;
        }
    }
    // remove_point (:342-355) for every FREE point behind a solid wall
    for (int w = 0; w < 6; ++w) {
        if (!solid[w]) continue;
        for (int k = 0; k < h.N; ++k) {
            if (!free_at(k) || !(wall_dist(k, w) < 0)) continue;
            const int pv = h.conn[2 * (size_t)k + 1], nx = h.conn[2 * (size_t)k];
            if (pv >= 0) h.conn[2 * (size_t)pv] = nx;
            if (nx >= 0) h.conn[2 * (size_t)nx + 1] = pv;
            h.conn[2 * (size_t)k] = h.conn[2 * (size_t)k + 1] = -1;
            h.st[2 * (size_t)k] = OVG_EMPTY;
            for (int sl : {k, pv, nx})
                if (sl >= 0) touched.push_back(sl);
            ++removed;
        }
    }
    c.next_empty = h.next_empty;
    *n_clip = clips;
    *n_rem = removed;
    if (clips || removed) {
        if (h.grown) {
            OVG_TRY(grow_slots(c, h.N, h.X));
        } else {
            for (int sl : touched) mark_dirty(c, sl);
            for (int nw : inserted) {
                if (c.h_status[2 * (size_t)nw] == OVG_EMPTY) continue;
                c.new_pos.push_back(nw);
                for (int q = 0; q < 3; ++q) c.new_xyz.push_back(h.X[3 * (size_t)nw + q]);
            }
            OVG_TRY(commit_topology(c));
        }
        OVG_TRY(refresh_pins(c));
    }
    // the state update_velocity left at every visited point
    const int nv = (int)visited.size();
    if (nv) {
        OVG_TRY(c.topo_stage.ensure(sizeof(int) * (size_t)nv + sizeof(double) * 12 * (size_t)nv +
                                    256));
        unsigned char *d = c.topo_stage.as<unsigned char>();
        const size_t off = ((sizeof(int) * (size_t)nv) + 255) & ~(size_t)255;
        OVG_HIP(hipMemcpyAsync(d, visited.data(), sizeof(int) * (size_t)nv, hipMemcpyHostToDevice,
                               c.stream));
        OVG_HIP(hipMemcpyAsync(d + off, state.data(), sizeof(double) * 12 * (size_t)nv,
                               hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(k_scatter_state, dim3((nv + 255) / 256), dim3(256), 0, c.stream, nv,
                           reinterpret_cast<const int *>(d),
                           reinterpret_cast<const double *>(d + off), dp(c.tan[0]), dp(c.tan[1]),
                           dp(c.tan[2]), dp(c.nor[0]), dp(c.nor[1]), dp(c.nor[2]), dp(c.vs[0]),
                           dp(c.vs[1]), dp(c.vs[2]), dp(c.vels[0]), dp(c.vels[1]),
                           dp(c.vels[2]));
        OVG_HIP(hipStreamSynchronize(c.stream));
    }
    return launch_status();
}

extern "C" int ovg_enforce_walls(ovg_ctx *ctx, double t, int *n_clipped, int *n_removed)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_mirrors(ctx->c));
    OVG_TRY(single_tangle(ctx->c, __func__));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    int a = 0, b = 0;
    const int r = walls_pass(c, t, &a, &b);
    if (n_clipped) *n_clipped = a;
    if (n_removed) *n_removed = b;
    return r;
}

// The reference visits every slot in index order (src/tangle.c:1021).  A
// slot whose links are unchanged when the loop reaches it decides exactly as
// it would with the links before the pass (positions of existing points do
// not move), so the pass visits, in index order, the device's candidates
// (that test with the links before the pass, widened by a margin) and every
// slot its own edits reach ahead of the cursor; all other slots would do
// nothing.
extern "C" int ovg_remesh(ovg_ctx *ctx, double min_dist, double max_dist, int *n_added,
                          int *n_removed)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(single_tangle(ctx->c, __func__));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    if (n_added) *n_added = 0;
    if (n_removed) *n_removed = 0;
    if (devtopo_applies(c)) {
        OVG_TRY(devtopo_remesh(c, min_dist, max_dist));
        if (n_added || n_removed) {
            int cnt[4];
            OVG_TRY(devtopo_counts(c, cnt));
            if (n_added) *n_added = cnt[2];
            if (n_removed) *n_removed = cnt[3];
        }
        return launch_status();
    }
    OVG_TRY(sync_mirrors(c));
    std::vector<int> cands;
    OVG_TRY(remesh_candidates(c, min_dist, max_dist, cands));
    if (cands.empty()) return OVG_OK;
    const double *Xp = nullptr;
    OVG_TRY(host_positions(c, &Xp));
    HostTangle h(c.n, Xp, c.h_conn, c.h_status, c.next_empty);
    TangleParams P = host_tparams(c.params, 0.0);
    int added = 0, removed = 0;
    std::priority_queue<int, std::vector<int>, std::greater<int>> ahead;
    int cursor = -1;
    size_t ci = 0;
    std::vector<int> touched;
    auto touch = [&](int sl) {
        if (sl < 0) return;
        touched.push_back(sl);
        if (sl > cursor) ahead.push(sl);
    };
    std::vector<int> inserted;
    for (;;) {
        const int a = ci < cands.size() ? cands[ci] : INT32_MAX;
        const int b = ahead.empty() ? INT32_MAX : ahead.top();
        const int k = a < b ? a : b;
        if (k == INT32_MAX) break;
        if (a == k) ++ci;
        while (!ahead.empty() && ahead.top() == k) ahead.pop();
        if (k <= cursor) continue;
        cursor = k;
        if (k >= h.N || h.st[2 * (size_t)k] == OVG_EMPTY) continue;
        int next = h.conn[2 * (size_t)k], prev = h.conn[2 * (size_t)k + 1];
        double lf = 0, lr = 0;
        const double *xk = &h.X[3 * (size_t)k];
        if (next >= 0) {
            double w[3] = {h.X[3 * (size_t)next], h.X[3 * (size_t)next + 1],
                           h.X[3 * (size_t)next + 2]};
            host::pwrap(P, xk, w);
            lf = host::dist3(xk, w);
        }
        if (prev >= 0) {
            double w[3] = {h.X[3 * (size_t)prev], h.X[3 * (size_t)prev + 1],
                           h.X[3 * (size_t)prev + 2]};
            host::pwrap(P, xk, w);
            lr = host::dist3(xk, w);
        }
        if (next >= 0 && prev >= 0 && ((lf < min_dist || lr < min_dist) && (lf + lr) < max_dist)) {
            int pv = h.conn[2 * (size_t)k + 1], nx = h.conn[2 * (size_t)k];  // remove_point :342-355
            if (pv >= 0) h.conn[2 * (size_t)pv] = nx;
            if (nx >= 0) h.conn[2 * (size_t)nx + 1] = pv;
            h.conn[2 * (size_t)k] = h.conn[2 * (size_t)k + 1] = -1;
            h.st[2 * (size_t)k] = OVG_EMPTY;
            ++removed;
            touch(k);
            touch(pv);
            touch(nx);
        }
        if (next >= 0 && lf > max_dist) {
            ++added;
            const int nxt = h.conn[2 * (size_t)k];
            const int nw = add_point(h, P, k, c.has_pinned ? h.st.data() : nullptr);
            inserted.push_back(nw);
            touch(nw);
            touch(k);
            touch(nxt);
        }
    }
    c.next_empty = h.next_empty;
    if (n_added) *n_added = added;
    if (n_removed) *n_removed = removed;
    if (!added && !removed) return OVG_OK;
    if (h.grown) {
        // the tangle doubled: reallocate and upload the whole state
        OVG_TRY(grow_slots(c, h.N, h.X));
    } else {
        for (int sl : touched) mark_dirty(c, sl);
        for (int nw : inserted) {
            if (c.h_status[2 * (size_t)nw] == OVG_EMPTY) continue;
            c.new_pos.push_back(nw);
            for (int q = 0; q < 3; ++q) c.new_xyz.push_back(h.X[3 * (size_t)nw + q]);
        }
        OVG_TRY(commit_topology(c));
    }
    if (added) {
        // enforce_periodic_boundaries and enforce_wall_boundaries after
        // insertions (:1057-1060), the latter at time 0 as the reference
        // driver calls remesh (src/main.c:131)
        launch_wrap(c);
        OVG_HIP(hipStreamSynchronize(c.stream));
        int clipped = 0, behind = 0;
        OVG_TRY(walls_pass(c, 0.0, &clipped, &behind));
    }
    return launch_status();
}

// add_loop, src/vortex_injection.c:44-103: the loop's points (the geometry
// of ovg_loop_points) in the tangle's next empty slots, linked in order and
// closed; the tangle doubles first while it has fewer empty slots than
// points (:50).  The new links and points reach the device in one commit.
extern "C" int ovg_add_loop(ovg_ctx *ctx, const double center[3], const double dir[3], double r,
                            double dl_min, double dl_max, int *n_points)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_mirrors(ctx->c));
    OVG_TRY(single_tangle(ctx->c, __func__));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    if (n_points) *n_points = 0;
    int npts = 0;
    if (ovg_loop_points(center, dir, r, dl_min, dl_max, nullptr, 0, &npts) != OVG_OK)
        return set_error(OVG_EINVAL, "add_loop: bad arguments");
    if (npts < 1)
        return set_error(OVG_EINVAL, "add_loop: radius %g gives %d points with dl %g..%g", r, npts,
                         dl_min, dl_max);
    std::vector<double> pts(3 * (size_t)npts);
    ovg_loop_points(center, dir, r, dl_min, dl_max, pts.data(), npts, &npts);
    const double *Xp = nullptr;
    OVG_TRY(host_positions(c, &Xp));
    HostTangle h(c.n, Xp, c.h_conn, c.h_status, c.next_empty);
    while (h.empty_count() < npts) h.expand();
    std::vector<int> slots;
    int first = -1, last = -1, cur = -1;
    for (int k = 0; k < npts; ++k) {  // :73-97
        cur = h.next_empty_node();
        if (first < 0) first = cur;
        h.st[2 * (size_t)cur] = OVG_FREE;
        for (int q = 0; q < 3; ++q) h.X[3 * (size_t)cur + q] = pts[3 * (size_t)k + q];
        if (cur != first) {
            h.conn[2 * (size_t)cur + 1] = last;
            h.conn[2 * (size_t)last] = cur;
        }
        last = cur;
        slots.push_back(cur);
    }
    h.conn[2 * (size_t)cur] = first;  // :100-101
    h.conn[2 * (size_t)first + 1] = cur;
    c.next_empty = h.next_empty;
    if (h.grown) {
        OVG_TRY(grow_slots(c, h.N, h.X));
    } else {
        for (int sl : slots) {
            mark_dirty(c, sl);
            c.new_pos.push_back(sl);
            for (int q = 0; q < 3; ++q) c.new_xyz.push_back(h.X[3 * (size_t)sl + q]);
        }
        OVG_TRY(commit_topology(c));
    }
    if (n_points) *n_points = npts;
    return launch_status();
}

// add_line_in_zdir, src/vortex_injection.c:267-318: a straight line at
// (x, y) from the DOWN wall to the UP wall (direction > 0) or back, of
// (int)(2 (zmax - zmin) / (dl_max - dl_min)) points (the reference divides
// by the difference here, by the sum in add_line), its ends of status
// pin_mode pinned to the walls; the tangle doubles first while it has fewer
// empty slots than points.
extern "C" int ovg_add_line_z(ovg_ctx *ctx, double x, double y, int direction, double dl_min,
                              double dl_max, int *n_points)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_mirrors(ctx->c));
    OVG_TRY(single_tangle(ctx->c, __func__));
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    if (n_points) *n_points = 0;
    const double zmin = c.params.box_lo[2], zmax = c.params.box_hi[2];
    if (direction == 0 || !(dl_max > dl_min))
        return set_error(OVG_EINVAL, "add_line_z: direction must be +-1 and dl_max > dl_min");
    const int npts = (int)(2 * (zmax - zmin) / (dl_max - dl_min));
    if (npts < 2)
        return set_error(OVG_EINVAL, "add_line_z: %d points for a box of height %g", npts,
                         zmax - zmin);
    const double *Xp = nullptr;
    OVG_TRY(host_positions(c, &Xp));
    HostTangle h(c.n, Xp, c.h_conn, c.h_status, c.next_empty);
    while (h.empty_count() < npts) h.expand();
    const double dz = (zmax - zmin) / npts;
    const double zstart = direction > 0 ? zmin : zmax, zend = direction > 0 ? zmax : zmin;
    const int pin = c.params.pin_mode ? c.params.pin_mode : OVG_PINNED;
    std::vector<int> slots;
    auto place = [&](int sl, double z) {
        h.X[3 * (size_t)sl] = x;
        h.X[3 * (size_t)sl + 1] = y;
        h.X[3 * (size_t)sl + 2] = z;
        slots.push_back(sl);
    };
    int cur = h.next_empty_node();  // the first point, pinned
    place(cur, zstart);
    h.st[2 * (size_t)cur] = pin;
    h.st[2 * (size_t)cur + 1] = direction > 0 ? OVG_FACE_DOWN : OVG_FACE_UP;
    h.conn[2 * (size_t)cur + 1] = -1;
    for (int k = 1; k < npts - 1; ++k) {
        const int last = cur;
        cur = h.next_empty_node();
        place(cur, zstart + direction * k * dz);
        h.st[2 * (size_t)cur] = OVG_FREE;
        h.st[2 * (size_t)cur + 1] = -1;
        h.conn[2 * (size_t)cur + 1] = last;
        h.conn[2 * (size_t)last] = cur;
    }
    const int last = cur;  // the last point, pinned
    cur = h.next_empty_node();
    place(cur, zend);
    h.st[2 * (size_t)cur] = pin;
    h.st[2 * (size_t)cur + 1] = direction > 0 ? OVG_FACE_UP : OVG_FACE_DOWN;
    h.conn[2 * (size_t)cur + 1] = last;
    h.conn[2 * (size_t)cur] = -1;
    h.conn[2 * (size_t)last] = cur;
    c.next_empty = h.next_empty;
    if (h.grown) {
        OVG_TRY(grow_slots(c, h.N, h.X));
    } else {
        for (int sl : slots) {
            mark_dirty(c, sl);
            c.new_pos.push_back(sl);
            for (int q = 0; q < 3; ++q) c.new_xyz.push_back(h.X[3 * (size_t)sl + q]);
        }
        OVG_TRY(commit_topology(c));
    }
    OVG_TRY(refresh_pins(c));
    if (n_points) *n_points = npts;
    return launch_status();
}

// ---- autotuning of the list traversal --------------------------------------
extern "C" int ovg_autotune(ovg_ctx *ctx, int repeats, ovg_tune *out)
{
    clear_error();
    OVG_TRY(check_ready(ctx));
    OVG_TRY(sync_live(ctx->c));
    if (!out || repeats < 1) return set_error(OVG_EINVAL, "bad arguments to ovg_autotune");
    Ctx &c = ctx->c;
    if (c.params.method != OVG_METHOD_BH || c.params.traverse != OVG_TRAVERSE_LIST)
        return set_error(OVG_EINVAL, "autotuning needs the Barnes-Hut list traversal");
    hipSetDevice(c.device);
    std::memset(out, 0, sizeof *out);
    const ovg_params saved = c.params;
    ovg_params auto_level = c.params;
    auto_level.group_level = 0;
    const int L = list_group_level(auto_level);
    push_params(c, 0.0);
    // the interaction counters are restored afterwards: tuning is not work
    unsigned long long keep[2] = {0, 0};
    OVG_HIP(hipMemcpyAsync(keep, c.counter64.p, 16, hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    Triple x = tri(c.x);
    OVG_TRY(build_tree(c, x.v));
    hipEvent_t e0 = take_event(c), e1 = take_event(c);
    double best = 1e300;
    int n = 0;
    // one setting: a warm-up launch, then `repeats` timed ones
    auto time_row = [&](const ovg_tune_row &set) -> int {
        c.params.points_per_wave = set.points_per_wave;
        c.params.leaf_size = set.leaf_size;
        c.params.group_level = set.group_level;
        c.params.list_stack = set.list_stack;
        c.params.list_tile = set.list_tile;
        c.params.list_blocks = set.list_blocks;
        push_params(c, 0.0);
        OVG_TRY(launch_bh_list(c, x.v));
        unsigned long long cnt0[2] = {0, 0}, cnt1[2] = {0, 0};
        OVG_HIP(hipMemcpyAsync(cnt0, c.counter64.p, 16, hipMemcpyDeviceToHost, c.stream));
        OVG_HIP(hipEventRecord(e0, c.stream));
        for (int r = 0; r < repeats; ++r) OVG_TRY(launch_bh_list(c, x.v));
        OVG_HIP(hipEventRecord(e1, c.stream));
        OVG_HIP(hipMemcpyAsync(cnt1, c.counter64.p, 16, hipMemcpyDeviceToHost, c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        ovg_tune_row &row = out->rows[n];
        row = set;
        row.ms = ms / repeats;
        row.pairs = ((double)(cnt1[0] - cnt0[0]) + (double)(cnt1[1] - cnt0[1])) / repeats;
        if (row.ms < best) {
            best = row.ms;
            out->best = n;
        }
        ++n;
        return OVG_OK;
    };
    // stage one: group size, bucket size and group level at the given shape
    const int ppws[5] = {8, 16, 21, 32, 64}, leaves[4] = {1, 2, 4, 8};
    for (int pw : ppws)
        for (int lf : leaves)
            for (int dl = -1; dl <= 1; ++dl) {
                const int lvl = L + dl;
                if (lvl < 1 || lvl > 21) continue;
                ovg_tune_row set{};
                set.points_per_wave = pw;
                set.leaf_size = lf;
                set.group_level = lvl;
                set.list_stack = saved.list_stack;
                set.list_tile = saved.list_tile;
                set.list_blocks = saved.list_blocks;
                OVG_TRY(time_row(set));
            }
    // stage two: the launch shape at the best of stage one
    const ovg_tune_row first = out->rows[out->best];
    const int stacks[3] = {128, 256, 384}, tiles[2] = {96, 128}, grids[3] = {0, 8, 16};
    for (int st : stacks)
        for (int tl : tiles)
            for (int bl : grids) {
                ovg_tune_row set = first;
                set.list_stack = st;
                set.list_tile = tl;
                set.list_blocks = bl;
                if (n < OVG_TUNE_MAX_ROWS) OVG_TRY(time_row(set));
            }
    out->n_rows = n;
    c.ev_pool.push_back(e0);
    c.ev_pool.push_back(e1);
    c.params = saved;
    const ovg_tune_row &win = out->rows[out->best];
    c.params.points_per_wave = win.points_per_wave;
    c.params.leaf_size = win.leaf_size;
    c.params.group_level = win.group_level;
    c.params.list_stack = win.list_stack;
    c.params.list_tile = win.list_tile;
    c.params.list_blocks = win.list_blocks;
    push_params(c, 0.0);
    OVG_HIP(hipMemcpyAsync(c.counter64.p, keep, 16, hipMemcpyHostToDevice, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    return launch_status();
}
