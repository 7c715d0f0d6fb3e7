// Internal declarations shared by the host-side engine and the kernel
// translation units.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/openvort_gpu.h"
#include "device_common.h"

namespace ovg {

int set_error(int code, const char *fmt, ...);
// Clears the calling thread's error text and any pending HIP launch error.
void clear_error();
// r when r != OVG_OK; otherwise OVG_EDEVICE when a kernel launch since the
// last clear_error() failed (hipGetLastError), else OVG_OK.
int launch_status(int r = OVG_OK);

#define OVG_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::ovg::set_error(OVG_EDEVICE, "%s failed: %s (%s:%d)", #call,        \
                                    hipGetErrorString(e_), __FILE__, __LINE__);        \
    } while (0)

#define OVG_TRY(call)                 \
    do {                              \
        int r_ = (call);              \
        if (r_ != OVG_OK) return r_;  \
    } while (0)

// Grow-only device buffer.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t n);
    void release();
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// Octree node in DFS pre-order, 64 bytes so a wavefront-uniform node loads
// with one scalar 64-byte load.
struct Node {
    double cx, cy, cz;   // centre of mass  (src/octree.c:264-269)
    double qx, qy, qz;   // total circulation, sum of pwrapped ds (:272-282)
    double Lm;           // largest side of the cell's box (:303-306)
    int32_t count;       // points in the cell (N)
    int32_t skip;        // DFS index of the next node outside this subtree
};
static_assert(sizeof(Node) == 64, "Node must be 64 bytes");

// Per-node build bookkeeping (separate from the 64-byte traversal record).
struct NodeMeta {
    int32_t start;       // first sorted point
    int32_t level;       // cell level (0 = root box)
    int32_t pad0, pad1;
};

struct TreeBufs {
    DBuf keys[2], vals[2];   // radix sort ping-pong (uint64 key, int32 point)
    DBuf hist;               // radix histograms
    DBuf delta;              // int8 LCP levels of adjacent sorted keys
    DBuf counts, base;       // nodes starting at each sorted point, exclusive scan
    DBuf nodes, meta;        // Node / NodeMeta
    DBuf level_hist, level_list;
    DBuf scan_tmp;
    DBuf counters;           // int32: [0] n_in, [1] n_nodes, [2] fallback count,
                             //        [3] max level, [4..] per-level counts
    DBuf sorted_pts;         // sorted point (slot) ids, in-box first then dropped
    DBuf sx, sy, sz;         // FP64 positions in sorted order
    DBuf sqx, sqy, sqz;      // FP64 pwrapped segment vectors in sorted order
    DBuf grp_a, grp_b, grp_c;  // target-group construction scratch (int, n + 1)
    DBuf groups;             // int2 (first sorted point, count) per target group
    DBuf rank;               // sorted position of each live slot
    DBuf sub_tgt, sub_groups, sub_count;  // list traversal of a target subset
    DBuf rec, rec_count;     // recorded interaction lists (reuse_lists)
    DBuf fix, fix_tiles;     // tuned trees: fixed-point prefix sums of the sorted points
    DBuf roots;              // DFS index of each tangle's root node
};
// counters[kGroupCounter]: number of target groups of the list traversal
constexpr int kGroupCounter = 40;
// list traversal launch modes: walk; walk and record each group's accepted
// entries; replay the recorded entries at new positions (same tree); walk
// with mirror images (kMirror, chosen by the launch when an image reflects)
constexpr int kWalk = 0, kRecord = 1, kReplay = 2, kMirror = 3;
// evaluation phases (eval_bs)
constexpr int kEvalPlain = 0, kEvalStep = 1, kEvalStage = 2;

struct Ctx {
    int device = 0;
    int n_cu = 256;  // compute units of the device (persistent list grids)
    hipStream_t stream = nullptr;
    ovg_params params{};
    bool have_params = false;
    int n = 0;            // slots
    int n_live = 0;       // non-EMPTY slots
    int n_src = 0;        // non-EMPTY slots with a forward neighbour
    int timing = 0;
    // Forest: n_tangles > 1 tangles in consecutive slot ranges, their
    // parameters in a device table (tparams[t]) and a per-slot tangle index
    // (tid); only the evaluation and the RK4 step run on a forest.
    int n_tangles = 1;
    std::vector<ovg_params> forest_params;
    std::vector<int> forest_off;     // slot offsets, n_tangles + 1
    std::vector<int> forest_live;    // live points per tangle
    int forest_leaf = -1, forest_glevel = -1;  // launch settings the table holds
    DBuf tid;
    // HIP graph of a captured call sequence (ovg_graph_*)
    bool capturing = false;
    hipGraphExec_t graph_exec = nullptr;

    // state, SoA over slots
    DBuf x[3], vels[3], vs[3], tan[3], nor[3];
    DBuf fwd, rev, status;
    DBuf pinw;               // pin_wall per slot (-1 when not pinned); set while has_pinned
    bool has_pinned = false; // the resident tangle has PINNED / PINNED_SLIP nodes
    DBuf live, src;                 // slot lists
    // RK4 scratch
    DBuf xs[2][3], move[3], k[3], stan[3], snor[3];
    // Biot-Savart partials
    DBuf bs_prim[3], bs_img[3], partial, dpart;  // partial: host staging; dpart: direct chunks
    DBuf split_part;                // per-(copy, target) sums of scattered-target walks
    DBuf seg[6];                    // direct-sum sources: r1 xyz, pwrapped r2 xyz
    DBuf seg_ids;                   // int2 (slot, fwd) per source
    DBuf tparams;                   // TangleParams[1]
    DBuf flags;                     // int32 device flags ([0] error bits)
    DBuf counter64;                 // uint64 interaction counters
    TreeBufs tree;
    uint64_t tree_id = 0;           // bumped whenever the tree buffers are rebuilt or reused
    int list_mode = 0;              // next list launch: walk / record / replay
    bool rec_ok = false;            // tree.rec holds the lists of tree rec_tree
    uint64_t rec_tree = 0;
    int rec_n = 0;                  // live points when they were recorded
    int rec_groups = 0;
    // reconnection / smoothing scratch
    DBuf rc_cell, rc_sorted, rc_start, rc_pairs, rc_count, rc_tmp;
    DBuf rc_room;                       // device remesh: its insertion bound (one u64)
    int rc_cap = 0;                 // int2 pairs that rc_pairs holds
    DBuf sm_buf;

    std::vector<int> h_conn, h_status;  // host mirrors of topology
    uint64_t topo_version = 1;          // bumped whenever links change
    // Incremental topology: host passes record the slots they touch; a
    // commit scatters just those slots to the device and rebuilds the live
    // list there (k_topo.hip).
    std::vector<int> dirty;             // slots with changed links/status
    std::vector<char> dirty_mark;
    std::vector<int> new_pos;           // slots with new positions (inserted points)
    std::vector<double> new_xyz;        // their positions, 3 per slot
    std::vector<int> loop_dirty;        // slots whose loop may have shrunk or split
    std::vector<char> loop_dirty_mark;
    bool loops_full_scan = true;        // next small-loop pass visits every loop
    int small_cutoff_seen = 0;          // largest cutoff of a completed pass
    double *h_xpin = nullptr;           // pinned AoS position mirror
    size_t h_xpin_cap = 0;              // doubles
    int xpin_state = 0;                 // mirror: 0 stale, 1 current, 2 current unless
                                        // the wraps since moved points (flags[1])
    bool xpin_pending = false;          // a copy into the mirror is in flight
    hipStream_t copy_stream = nullptr;  // position downloads beside the main stream
    hipEvent_t xpin_from = nullptr, xpin_done = nullptr;
    DBuf xaos;                          // AoS staging of those downloads
    std::vector<int> rc_where, rc_flag; // reconnection scratch, kept -1 / 0
    DBuf topo_stage;                    // device staging of commits
    int next_empty = 0;                 // the reference allocator's hint (tangle->next_empty)
    uint64_t loops_version = 0;         // topology the smoothing loop table matches
    int n_loops = 0;
    DBuf loop_off, loop_slots;          // smoothing loop table (CSR over slots)
    DBuf lt_buf;                        // loop table construction scratch
    std::vector<int> h_loop_off;
    std::vector<int> h_loop_open;       // per loop: 1 = an open line (ends on walls)
    DBuf loop_open;
    DBuf sm_fft, sm_big;                // transform arena and big-loop list + chunk maps
    DBuf sm_gstate;                     // per big loop state, per chunk wrap sums
    DBuf sm_tw;                         // e^{-2 pi i k / 2^sm_tw_log}, k < 2^(sm_tw_log-1)
    int sm_tw_log = 0;
    std::vector<char> sm_hmeta;         // host copy of the batched transform's tables,
    hipEvent_t sm_hmeta_done = nullptr; // recorded after their upload: rewritten only
                                        // once that upload has completed
    // Device-resident topology (ovg_params.topology == OVG_TOPO_DEVICE):
    // the reconnection, small-loop and remesh passes edit the device links,
    // status and positions in place and queue without waiting; the live
    // count and the host mirrors catch up when a call needs them
    // (sync_live, sync_mirrors).
    DBuf dt_state;                      // int counters of the device passes (k_devtopo.hip)
    DBuf dt_where;                      // int per slot: endpoint index, -1 between passes
    DBuf dt_mark;                       // int per slot, 0 between passes
    DBuf dt_pairs, dt_ends, dt_chg;     // reconnection pass scratch
    DBuf dt_cl;                         // its clusters: union-find parents, flags
    DBuf dt_grp;                        // endpoint target groups: flags, scans, cells
    DBuf dt_delta;                      // direct mode: field change of each change at each end
    DBuf prem_buf;                      // remesh by clusters: parents, owners, copies, table
    int prem_seen = 0;                  // most remesh candidates a device pass met
    bool dt_unread = false;             // device passes ran since the counters were read
    bool live_pending = false;          // the device's live list is newer than n_live
    bool mirrors_stale = false;         // h_conn / h_status / next_empty are older
    bool hint_on_device = false;        // the device holds the allocator hint
    int dt_cap = 0;                     // candidate pairs a device pass takes
    int dt_seen = 0;                    // most candidate pairs a device pass met
    ovg_stats stats{};
    hipEvent_t ev[8] = {};
    // per-evaluation timing events, resolved when statistics are read so
    // that timing adds no synchronisation to the timed work
    std::vector<hipEvent_t> markers;     // ovg_marker_record, created on first use
    std::vector<hipEvent_t> ev_pool;
    std::vector<hipEvent_t> ev_pending;  // triples: start, tree done, BS done
};

// ---- kernel launchers (one translation unit each) ------------------------
// All take the context stream; none synchronises.

// k_local.hip
void launch_stencil(Ctx &c, const double *const x[3], double *const tan[3],
                    double *const nor[3]);
void launch_prep_segments(Ctx &c, const double *const x[3]);
void launch_compose(Ctx &c, const double *const x[3], bool own_tangents, int stage,
                    double dt);
void launch_stage_positions(Ctx &c, const double *const kv[3], double h,
                            double *const out[3]);
void launch_wrap(Ctx &c);
// enforce_periodic_boundaries over every slot that is not EMPTY, queued
// only when the device int *d_gate is nonzero (no wait).
void launch_wrap_slots(Ctx &c, const int *d_gate);
void launch_line_length(Ctx &c, double *d_out);
void launch_check_finite(Ctx &c, const double *const x[3]);

// k_direct.hip
int launch_direct(Ctx &c, const double *const x[3], int n_tgt, const int *d_tgt_slots,
                  double *const out_prim[3], double *const out_img[3], int chunks);

// The same sum with its chunk partials in the device's static scratch array
// (no allocation): d_slots is checked into scratch first (a slot out of
// range or EMPTY becomes -1 and sets bit 4 of the device flags); *checked
// receives that list.  At most OVG_NOALLOC_MAX_TARGETS targets.
int launch_direct_noalloc(Ctx &c, const double *const x[3], int n_tgt, const int *d_slots,
                          double *const out_prim[3], double *const out_img[3],
                          const int **checked);
// Resolve the device address of that scratch array for `device` (current).
int direct_scratch_init(int device);

// engine.hip: write a forest's parameter table (when the launch settings
// changed since the last write, or always with force)
int forest_table(Ctx &c, bool force);
// k_tree.hip: fixed-point units of a tangle of n points, largest side Lmax
void fix_units(int n, double Lmax, double &scale, double &inv, double &lim);

// k_tree.hip
int build_tree(Ctx &c, const double *const x[3]);
// Same tree (keys, order, nodes, cell sizes) with positions, segment vectors
// and node sums recomputed at x: the stages of one RK4 step.
int refit_tree(Ctx &c, const double *const x[3]);

// k_bh.hip
int launch_bh(Ctx &c, const double *const x[3]);
void launch_bh_targets(Ctx &c, const double *const x[3], int n_tgt, const int *tlist);
// k_list.hip
int launch_bh_list(Ctx &c, const double *const x[3]);
int list_group_level(const ovg_params &p);    // the level that bounds a target group
int list_points_per_wave(const ovg_params &p);
// The list traversal for m live slots (device list) against the tree built
// at x: grouped by tree order within level-L0 cells.
int launch_bh_list_targets(Ctx &c, const double *const x[3], int m, const int *d_slots);
// The list traversal of target groups already built on the device: d_grp
// (first index into d_tgt, count) per group, their number at *d_ngroups,
// at most max_groups.
int launch_bh_list_groups(Ctx &c, const double *const x[3], const int2 *d_grp, const int *d_tgt,
                          const int *d_ngroups, int max_groups);

// k_topo.hip
// Pinned host copy of the positions (AoS, c.n slots), valid until the next
// device-side change of positions.
int host_positions(Ctx &c, const double **X);
// Start refreshing the pinned mirror on the copy stream behind the work
// queued so far (no-op when it is current); host_positions waits for it.
int host_positions_begin(Ctx &c);
// Call before queueing any device write of positions: waits for a copy in
// flight and marks the mirror stale.
int positions_change(Ctx &c);
// Record a slot whose links or status changed (also for the small-loop pass).
void mark_dirty(Ctx &c, int slot);
// Scatter the dirty slots' links/status and the new positions to the device,
// rebuild the live list there and bump topo_version.
int commit_topology(Ctx &c);
// Remesh candidates: live slots whose original-link decision may remove or
// insert (a superset of the slots the sequential pass acts on first).
int remesh_candidates(Ctx &c, double min_dist, double max_dist, std::vector<int> &out);
// The same candidates queued on the device: the ascending slot list at
// *d_list, its length at *d_count (no wait).  d_room, when given, gains a
// bound on the points the pass can insert.
int remesh_candidates_queue(Ctx &c, double min_dist, double max_dist, const int **d_list,
                            const int **d_count, unsigned long long *d_room = nullptr);

// k_recon.hip
int recon_candidates(Ctx &c, int max_pairs, int *h_pairs, int *n_pairs);
// Queue the candidate scan: pairs at c.rc_pairs (unordered, at most
// c.rc_cap of them), their full count at c.rc_count[0]; no wait.
int candidates_queue(Ctx &c);

// k_topo.hip: ordered live list over every slot, rebuilt on the device;
// the count goes to d_count (device int).  No wait.
int rebuild_live(Ctx &c, int *d_count);

// k_devtopo.hip: the device-resident topology passes
// (ovg_params.topology == OVG_TOPO_DEVICE)
bool devtopo_applies(const Ctx &c);   // the tangle qualifies (one periodic tangle, no pins)
int devtopo_reconnect(Ctx &c, double t);
int devtopo_small_loops(Ctx &c, int cutoff);
int devtopo_remesh(Ctx &c, double min_dist, double max_dist);
// The counters of the last device passes (waits): reconnections, small-loop
// points removed, remesh points added and removed.
int devtopo_counts(Ctx &c, int out[4]);
// The 20 public device counters (ovg_topology_counters).
int devtopo_counters(Ctx &c, int out[22]);
// engine.hip: grow the slots to n_new (> n) with the host mirrors
// current, the new slots EMPTY; every old slot keeps its state.
int grow_capacity(Ctx &c, int n_new);
// n_live / n_src from the device after device passes changed the live set
// (waits once); reports an error a device pass left.
int sync_live(Ctx &c);
// h_conn / h_status / next_empty from the device after device passes
// (waits); implies sync_live.
int sync_mirrors(Ctx &c);

// k_smooth.hip
int smooth_loops(Ctx &c);

// engine.hip helpers
// Biot-Savart (primary + images) of every live point at positions x with
// the context's method, into c.bs_prim / c.bs_img.
int eval_bs(Ctx &c, const double *const x[3], int phase);
TangleParams host_tparams(const ovg_params &p, double t);
void push_params(Ctx &c, double t);
void launch_soa_to_aos(Ctx &c, const double *x, const double *y, const double *z, double *aos);
void launch_soa_to_aos(Ctx &c, const double *x, const double *y, const double *z, double *aos,
                       hipStream_t stream);
void launch_split_pairs(Ctx &c, const int *pairs);

// k_recon.hip: candidate scan + greedy pass; links updated on the device
int reconnect_pass(Ctx &c, double t, int *n_rec);

// k_tree.hip: stable LSD radix sort of c.tree.keys[0]/vals[0] over the
// 8-bit digits at first_shift..last_shift (an even number of passes)
int radix_sort(Ctx &c, int n, int first_shift, int last_shift);

// scan helper (k_tree.hip): exclusive scan of n int32 values, in -> out,
// writes the total to *d_total when non-null.
// Bits below the octree cell of `level` in a tuned tree key (one tangle).
int tuned_key_cell_shift(int level);
int launch_exclusive_scan(Ctx &c, const int *d_in, int *d_out, int n, int *d_total,
                          DBuf &tmp);

inline const TangleParams *dparams(const Ctx &c) { return c.tparams.as<TangleParams>(); }

}  // namespace ovg
