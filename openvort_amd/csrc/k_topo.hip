// Incremental topology maintenance between the host passes that follow the
// reference's sequential semantics (remesh, reconnection, small-loop
// removal) and the device state.
//
// A pass reads a pinned host copy of the positions, edits the host mirrors
// of the links and status (Ctx::h_conn / h_status) and records every slot it
// touches.  commit_topology then moves only those slots to the device (a
// scatter of (slot, forward, reverse, status) records and of the inserted
// points' positions) and rebuilds the ordered live-slot list on the device
// with a flag / exclusive-scan / scatter pass.  remesh_candidates runs the
// remesh test (src/tangle.c:1039-1051) with the links as they are before
// the pass on every slot at once, so the host pass only visits candidates
// and the slots its own edits reach.
// HBM-bound, tiny: O(changed slots) transfers and O(slots) device passes.
#include <cstring>

#include "engine_internal.h"

namespace ovg {

// Pinned mirror large enough for the current slots (its contents are lost
// when it grows).
static int mirror_capacity(Ctx &c)
{
    const size_t need = 3 * (size_t)(c.n > 0 ? c.n : 1);
    if (need <= c.h_xpin_cap) return OVG_OK;
    if (c.h_xpin) hipHostFree(c.h_xpin);
    c.h_xpin = nullptr;
    c.h_xpin_cap = 0;
    c.xpin_state = 0;
    size_t cap = need + need / 4;
    if (hipHostMalloc((void **)&c.h_xpin, sizeof(double) * cap, hipHostMallocDefault) !=
        hipSuccess) {
        c.h_xpin = nullptr;
        return set_error(OVG_ENOMEM, "pinned host buffer of %zu doubles", cap);
    }
    c.h_xpin_cap = cap;
    return OVG_OK;
}

static int settle(Ctx &c)
{
    if (!c.xpin_pending) return OVG_OK;
    c.xpin_pending = false;
    OVG_HIP(hipEventSynchronize(c.xpin_done));
    c.xpin_state = 1;
    return OVG_OK;
}

int host_positions_begin(Ctx &c)
{
    if (c.xpin_pending || c.xpin_state == 1 || c.n <= 0) return OVG_OK;
    OVG_TRY(mirror_capacity(c));
    const size_t bytes = sizeof(double) * 3 * (size_t)c.n;
    OVG_TRY(c.xaos.ensure(bytes));
    OVG_HIP(hipEventRecord(c.xpin_from, c.stream));
    OVG_HIP(hipStreamWaitEvent(c.copy_stream, c.xpin_from, 0));
    launch_soa_to_aos(c, c.x[0].as<double>(), c.x[1].as<double>(), c.x[2].as<double>(),
                      c.xaos.as<double>(), c.copy_stream);
    OVG_HIP(hipMemcpyAsync(c.h_xpin, c.xaos.p, bytes, hipMemcpyDeviceToHost, c.copy_stream));
    OVG_HIP(hipEventRecord(c.xpin_done, c.copy_stream));
    c.xpin_pending = true;
    return OVG_OK;
}

int positions_change(Ctx &c)
{
    OVG_TRY(settle(c));
    c.xpin_state = 0;
    return OVG_OK;
}

int host_positions(Ctx &c, const double **X)
{
    OVG_TRY(settle(c));
    OVG_TRY(mirror_capacity(c));
    if (c.xpin_state == 2) {
        int moved = 0;
        OVG_HIP(hipMemcpyAsync(&moved, c.flags.as<int>() + 1, sizeof(int), hipMemcpyDeviceToHost,
                               c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
        c.xpin_state = moved ? 0 : 1;
    }
    if (c.xpin_state == 1) {
        *X = c.h_xpin;
        return OVG_OK;
    }
    if (c.n > 0) {
        OVG_TRY(c.partial.ensure(sizeof(double) * 3 * (size_t)c.n));
        launch_soa_to_aos(c, c.x[0].as<double>(), c.x[1].as<double>(), c.x[2].as<double>(),
                          c.partial.as<double>());
        OVG_HIP(hipMemcpyAsync(c.h_xpin, c.partial.p, sizeof(double) * 3 * (size_t)c.n,
                               hipMemcpyDeviceToHost, c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
    }
    c.xpin_state = 1;
    *X = c.h_xpin;
    return OVG_OK;
}

void mark_dirty(Ctx &c, int slot)
{
    if (slot < 0) return;
    if ((size_t)slot >= c.dirty_mark.size()) c.dirty_mark.resize((size_t)c.n, 0);
    if ((size_t)slot >= c.loop_dirty_mark.size()) c.loop_dirty_mark.resize((size_t)c.n, 0);
    if (!c.dirty_mark[slot]) {
        c.dirty_mark[slot] = 1;
        c.dirty.push_back(slot);
    }
    if (!c.loop_dirty_mark[slot]) {
        c.loop_dirty_mark[slot] = 1;
        c.loop_dirty.push_back(slot);
    }
}

// (slot, forward, reverse, status) records
__global__ void k_scatter_topo(int m, const int4 *__restrict__ rec, int *__restrict__ fwd,
                               int *__restrict__ rev, int *__restrict__ status)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int4 r = rec[i];
    fwd[r.x] = r.y;
    rev[r.x] = r.z;
    status[r.x] = r.w;
}

__global__ void k_scatter_pos(int m, const int *__restrict__ slot, const double *__restrict__ xyz,
                              double *__restrict__ x, double *__restrict__ y,
                              double *__restrict__ z)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int k = slot[i];
    x[k] = xyz[3 * i];
    y[k] = xyz[3 * i + 1];
    z[k] = xyz[3 * i + 2];
}

__global__ void k_live_flags(int n, const int *__restrict__ status, int *__restrict__ flag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = status[i] != OVG_EMPTY;
}

__global__ void k_zero_slots(int n, int *__restrict__ a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void k_live_scatter(int n, const int *__restrict__ status,
                               const int *__restrict__ pos, int *__restrict__ live,
                               int *__restrict__ src)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || status[i] == OVG_EMPTY) return;
    live[pos[i]] = i;
    src[pos[i]] = i;
}

int rebuild_live(Ctx &c, int *d_count)
{
    const int n = c.n;
    OVG_TRY(c.dt_mark.ensure(sizeof(int) * (size_t)(n + 1)));
    int *pos = c.dt_mark.as<int>();
    hipLaunchKernelGGL(k_live_flags, dim3((n + 255) / 256), dim3(256), 0, c.stream, n,
                       c.status.as<int>(), pos);
    OVG_TRY(launch_exclusive_scan(c, pos, pos, n, d_count, c.tree.scan_tmp));
    hipLaunchKernelGGL(k_live_scatter, dim3((n + 255) / 256), dim3(256), 0, c.stream, n,
                       c.status.as<int>(), pos, c.live.as<int>(), c.src.as<int>());
    // dt_mark is zero between passes
    hipLaunchKernelGGL(k_zero_slots, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, pos);
    return OVG_OK;
}

int commit_topology(Ctx &c)
{
    c.rec_ok = false;  // recorded interaction lists follow the old links
    const int m = (int)c.dirty.size(), a = (int)c.new_pos.size();
    if (m == 0 && a == 0) return OVG_OK;
    const int n = c.n;
    // host records, then one staging upload
    const size_t rec_bytes = sizeof(int4) * (size_t)m;
    const size_t pos_bytes = sizeof(int) * (size_t)a, xyz_bytes = sizeof(double) * 3 * (size_t)a;
    const size_t off_pos = (rec_bytes + 255) & ~(size_t)255;
    const size_t off_xyz = (off_pos + pos_bytes + 255) & ~(size_t)255;
    const size_t total = off_xyz + xyz_bytes;
    std::vector<unsigned char> h(total);
    int4 *rec = reinterpret_cast<int4 *>(h.data());
    int live = 0;
    for (int i = 0; i < m; ++i) {
        const int k = c.dirty[i];
        rec[i] = make_int4(k, c.h_conn[2 * (size_t)k], c.h_conn[2 * (size_t)k + 1],
                           c.h_status[2 * (size_t)k]);
        c.dirty_mark[k] = 0;
    }
    if (a) {
        std::memcpy(h.data() + off_pos, c.new_pos.data(), pos_bytes);
        std::memcpy(h.data() + off_xyz, c.new_xyz.data(), xyz_bytes);
    }
    OVG_TRY(c.topo_stage.ensure(total));
    OVG_HIP(hipMemcpyAsync(c.topo_stage.p, h.data(), total, hipMemcpyHostToDevice, c.stream));
    unsigned char *d = c.topo_stage.as<unsigned char>();
    if (m)
        hipLaunchKernelGGL(k_scatter_topo, dim3((m + 255) / 256), dim3(256), 0, c.stream, m,
                           reinterpret_cast<const int4 *>(d), c.fwd.as<int>(), c.rev.as<int>(),
                           c.status.as<int>());
    if (a)
        hipLaunchKernelGGL(k_scatter_pos, dim3((a + 255) / 256), dim3(256), 0, c.stream, a,
                           reinterpret_cast<const int *>(d + off_pos),
                           reinterpret_cast<const double *>(d + off_xyz), c.x[0].as<double>(),
                           c.x[1].as<double>(), c.x[2].as<double>());
    // ordered live list on the device
    OVG_TRY(c.rc_tmp.ensure(sizeof(int) * (size_t)(n + 1)));
    int *flag = c.rc_tmp.as<int>();
    OVG_TRY(rebuild_live(c, flag + n));
    OVG_HIP(hipMemcpyAsync(&live, flag + n, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    c.n_live = live;
    c.n_src = live;  // closed loops: every live slot starts a segment
    c.dirty.clear();
    c.new_pos.clear();
    c.new_xyz.clear();
    c.topo_version++;
    return OVG_OK;
}

// Remesh test of src/tangle.c:1039-1051 with the current links, widened by
// a relative margin so that rounding differences between this evaluation
// and the host pass can only add candidates.
__global__ void k_remesh_flags(int n, const int *__restrict__ status, const int *__restrict__ fwd,
                               const int *__restrict__ rev, const double *__restrict__ X,
                               const double *__restrict__ Y, const double *__restrict__ Z,
                               const TangleParams *__restrict__ Pp, double min_dist,
                               double max_dist, int *__restrict__ flag,
                               unsigned long long *__restrict__ room)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int f = 0;
    unsigned long long need = 0;
    if (status[k] != OVG_EMPTY) {
        const TangleParams &P = *Pp;
        const int nx = fwd[k], pv = rev[k];
        const double x = X[k], y = Y[k], z = Z[k];
        double lf = 0, lr = 0;
        if (nx >= 0) {
            double a = X[nx], b = Y[nx], cz = Z[nx];
            pwrap(P, x, y, z, a, b, cz);
            a -= x; b -= y; cz -= z;
            lf = sqrt(a * a + b * b + cz * cz);
        }
        if (pv >= 0) {
            double a = X[pv], b = Y[pv], cz = Z[pv];
            pwrap(P, x, y, z, a, b, cz);
            a -= x; b -= y; cz -= z;
            lr = sqrt(a * a + b * b + cz * cz);
        }
        const double up = 1.0 + 1e-9, dn = 1.0 - 1e-9;
        const bool rm = nx >= 0 && pv >= 0 && (lf < min_dist * up || lr < min_dist * up) &&
                        (lf + lr) < max_dist * up;
        const bool ad = nx >= 0 && lf > max_dist * dn;
        // a non-finite length decides nothing in the reference either way;
        // let the host look at it
        f = rm || ad || !(lf == lf) || !(lr == lr);
        // room: a bound on the points the pass can insert for this segment
        // (each split leaves pieces at most 1/sqrt(2) of the chord, so a
        // chain of splits of one segment is shorter than 2 lf / max_dist)
        if (ad) need = lf < 1e6 * max_dist ? (unsigned long long)(2.0 * lf / max_dist) + 1 : 1ull << 20;
        else if (f) need = 1;
    }
    flag[k] = f;
    if (room && need) atomicAdd(room, need);
}

__global__ void k_compact(int n, const int *__restrict__ flag, const int *__restrict__ pos,
                          int *__restrict__ out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && flag[k]) out[pos[k]] = k;
}

int remesh_candidates_queue(Ctx &c, double min_dist, double max_dist, const int **d_list,
                            const int **d_count, unsigned long long *d_room)
{
    const int n = c.n;
    OVG_TRY(c.rc_tmp.ensure(sizeof(int) * (size_t)(n + 1)));
    OVG_TRY(c.rc_count.ensure(sizeof(int) * (size_t)(n + 1)));
    OVG_TRY(c.rc_sorted.ensure(sizeof(int) * (size_t)(n + 1)));
    int *flag = c.rc_tmp.as<int>(), *pos = c.rc_count.as<int>(), *lst = c.rc_sorted.as<int>();
    push_params(c, 0.0);
    hipLaunchKernelGGL(k_remesh_flags, dim3((n + 255) / 256), dim3(256), 0, c.stream, n,
                       c.status.as<int>(), c.fwd.as<int>(), c.rev.as<int>(), c.x[0].as<double>(),
                       c.x[1].as<double>(), c.x[2].as<double>(), dparams(c), min_dist, max_dist,
                       flag, d_room);
    OVG_TRY(launch_exclusive_scan(c, flag, pos, n, pos + n, c.tree.scan_tmp));
    hipLaunchKernelGGL(k_compact, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, flag, pos,
                       lst);
    *d_list = lst;
    *d_count = pos + n;
    return OVG_OK;
}

int remesh_candidates(Ctx &c, double min_dist, double max_dist, std::vector<int> &out)
{
    const int n = c.n;
    out.clear();
    if (n == 0) return OVG_OK;
    const int *lst = nullptr, *cntp = nullptr;
    OVG_TRY(remesh_candidates_queue(c, min_dist, max_dist, &lst, &cntp));
    const int *pos = cntp - n;
    int cnt = 0;
    OVG_HIP(hipMemcpyAsync(&cnt, pos + n, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    OVG_HIP(hipStreamSynchronize(c.stream));
    out.resize(cnt);
    if (cnt) {
        OVG_HIP(hipMemcpyAsync(out.data(), lst, sizeof(int) * (size_t)cnt, hipMemcpyDeviceToHost,
                               c.stream));
        OVG_HIP(hipStreamSynchronize(c.stream));
    }
    return OVG_OK;
}

}  // namespace ovg
