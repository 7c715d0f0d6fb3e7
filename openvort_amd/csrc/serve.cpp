// Request loop: admission by estimated HBM footprint, device-resident
// state per job, concurrent stepping by worker threads (see
// include/openvort_serve.h).  Host code over the C ABI only.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/openvort_gpu.h"
#include "../../include/openvort_serve.h"

namespace {

using Clock = std::chrono::steady_clock;

double since(Clock::time_point a) { return std::chrono::duration<double>(Clock::now() - a).count(); }

struct Active {
    int job = -1;
    ovg_ctx *ctx = nullptr;
    int steps_done = 0;
    int n_points = 0;
    double t = 0;      // simulation time
    double bytes = 0;  // admitted footprint estimate
    bool busy = false;
    bool ready = false;
};

// The workload parameters of src/initial_conditions.c scaled to the job's
// point spacing: rec_dist = 0.8 dl_min, dl_min = dxi/2, dl_max = 1.5 dxi,
// reconnection angle 5 degrees, filter cutoff 0.8; kappa, a0 of helium II.
void job_params(const ovs_job &j, ovg_params &p)
{
    ovg_params_default(&p, j.D);
    p.kappa = 9.97e-4;
    p.vortex_width = 1.3e-8;
    p.use_mutual_friction = 1;
    p.alpha = j.alpha;
    p.alpha_p = j.alpha_p;
    if (j.v_ns != 0) {
        p.vn.type = OVG_FLOW_SIMPLE;
        p.vn.strength = j.v_ns;
        p.vn.direction[0] = 1.0;
        p.vn.direction[1] = 0.0;
        p.vn.direction[2] = 0.0;
    }
    p.method = OVG_METHOD_BH;
    p.precision = j.precision;
    p.theta = j.theta > 0 ? j.theta : 0.4;
    p.traverse = j.traverse;
    p.leaf_size = j.leaf_size > 0 ? j.leaf_size : 1;
    p.points_per_wave = j.points_per_wave;
    p.rec_dist = 0.8 * (j.dxi / 2);
    p.rec_angle = 5.0 * M_PI / 180.0;
    p.freq_to_cutoff = 0.8;
    p.mixed_near = 2 * j.dxi;
    p.rec_velocity = j.rec_velocity;
    p.topology = j.topology;
}

int create_job(const ovs_job &j, Active &a)
{
    int n = ovg_generate_rings(j.seed, j.n_rings, j.D, j.r_min, j.r_max, j.dxi, 0, nullptr,
                               nullptr, nullptr);
    if (n <= 0) return OVG_EINVAL;
    int cap = 512;
    while (cap < n) cap *= 2;  // create_tangle sizes, src/main.c:85, src/tangle.c:248
    std::vector<double> x(3 * (size_t)cap);
    std::vector<int> conn(2 * (size_t)cap), st(2 * (size_t)cap);
    if (ovg_generate_rings(j.seed, j.n_rings, j.D, j.r_min, j.r_max, j.dxi, cap, x.data(),
                           conn.data(), st.data()) != n)
        return OVG_EINVAL;
    ovg_params p;
    job_params(j, p);
    int r = ovg_create(&a.ctx, 0);
    if (r) return r;
    if ((r = ovg_set_params(a.ctx, &p))) return r;
    if ((r = ovg_upload(a.ctx, cap, x.data(), conn.data(), st.data()))) return r;
    a.n_points = n;
    return OVG_OK;
}

// One step in the order of src/main.c:127-138.
// The passes get no count pointers, so that with device topology they queue
// without waiting for the host.
int run_step(const ovs_job &j, Active &a)
{
    int r;
    const int k = a.steps_done;
    if ((r = ovg_reconnect(a.ctx, a.t, nullptr))) return r;
    if ((r = ovg_eliminate_small_loops(a.ctx, 5, nullptr))) return r;
    if (j.smooth_every > 0 && k % j.smooth_every == 0 && (r = ovg_smooth(a.ctx))) return r;
    if ((r = ovg_enforce_periodic(a.ctx))) return r;
    if ((r = ovg_remesh(a.ctx, j.dxi / 2, 1.5 * j.dxi, nullptr, nullptr))) return r;
    if ((r = ovg_eliminate_small_loops(a.ctx, 5, nullptr))) return r;
    if ((r = ovg_update_tangle(a.ctx, a.t))) return r;
    if ((r = ovg_rk4_step(a.ctx, a.t, j.dt))) return r;
    if ((r = ovg_enforce_periodic(a.ctx))) return r;
    if ((r = ovg_synchronize(a.ctx))) return r;
    a.t += j.dt;
    a.n_points = ovg_num_points(a.ctx);
    return OVG_OK;
}

uint64_t fnv(uint64_t h, const void *p, size_t n)
{
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ULL;
    }
    return h;
}

void finish_job(Active &a, ovs_result *res)
{
    if (!a.ctx) return;
    if (res && !res->failed) {
        const int n = ovg_num_slots(a.ctx);
        std::vector<double> x(3 * (size_t)n);
        std::vector<int> conn(2 * (size_t)n), st(2 * (size_t)n);
        if (ovg_download(a.ctx, x.data(), nullptr, nullptr, nullptr, nullptr) == OVG_OK &&
            ovg_download_connections(a.ctx, conn.data(), st.data()) == OVG_OK) {
            uint64_t h = 1469598103934665603ULL;
            for (int k = 0; k < n; ++k) {
                if (st[2 * (size_t)k] == OVG_EMPTY) continue;
                h = fnv(h, &k, sizeof k);
                h = fnv(h, &x[3 * (size_t)k], 3 * sizeof(double));
                h = fnv(h, &conn[2 * (size_t)k], 2 * sizeof(int));
            }
            res->checksum = h;
        }
        double L = 0;
        int np = 0;
        if (ovg_line_length(a.ctx, &L, &np) == OVG_OK) {
            res->line_length = L;
            res->points_final = np;
        }
    }
    ovg_destroy(a.ctx);
    a.ctx = nullptr;
}

struct Server {
    const ovs_job *jobs;
    int n_jobs, max_active;
    double cap;
    bool realtime;
    ovs_result *res;
    Clock::time_point t0;

    std::mutex m;
    std::condition_variable cv;
    int next = 0;   // next job in arrival order
    int closed = 0; // finished or failed
    size_t rr = 0;  // round-robin cursor
    std::vector<Active *> active;
    double bytes = 0, peak_bytes = 0;
    int peak_active = 0;
    std::vector<double> step_ms;
    double point_steps = 0;

    bool arrived(int j) const { return !realtime || since(t0) >= jobs[j].arrival_s; }

    void close_job(std::unique_lock<std::mutex> &lk, Active *a)
    {
        active.erase(std::find(active.begin(), active.end(), a));
        bytes -= a->bytes;
        lk.unlock();
        finish_job(*a, res ? &res[a->job] : nullptr);
        lk.lock();
        if (res) {
            ovs_result &r = res[a->job];
            r.steps_done = a->steps_done;
            r.done_s = since(t0);
            r.latency_s = r.done_s - (realtime ? jobs[a->job].arrival_s : 0.0);
        }
        delete a;
        ++closed;
        cv.notify_all();
    }

    void worker()
    {
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            if (closed == n_jobs) {
                cv.notify_all();
                return;
            }
            // admission, in arrival order, while the footprint fits
            if (next < n_jobs && arrived(next) && (int)active.size() < max_active) {
                const int j = next;
                int n = ovg_generate_rings(jobs[j].seed, jobs[j].n_rings, jobs[j].D,
                                           jobs[j].r_min, jobs[j].r_max, jobs[j].dxi, 0,
                                           nullptr, nullptr, nullptr);
                int slots = 512;
                while (slots < n) slots *= 2;
                const double b = ovs_job_bytes(slots);
                if (bytes + b <= cap || (active.empty() && b <= cap)) {
                    ++next;
                    Active *a = new Active;
                    a->job = j;
                    a->bytes = b;
                    a->busy = true;
                    active.push_back(a);
                    bytes += b;
                    peak_bytes = std::max(peak_bytes, bytes);
                    peak_active = std::max(peak_active, (int)active.size());
                    lk.unlock();
                    const int r = create_job(jobs[j], *a);
                    lk.lock();
                    if (res) {
                        res[j].admit_s = since(t0);
                        res[j].points_initial = a->n_points;
                        res[j].failed = r;
                    }
                    if (r) {
                        close_job(lk, a);
                    } else {
                        a->busy = false;
                        a->ready = true;
                        cv.notify_all();
                    }
                    continue;
                }
                if (active.empty()) {  // larger than the whole cap
                    ++next;
                    if (res) res[j].failed = OVG_ENOMEM;
                    ++closed;
                    continue;
                }
            }
            // one step of a job nobody is stepping
            Active *pick = nullptr;
            for (size_t q = 0; q < active.size(); ++q) {
                Active *a = active[(rr + q) % active.size()];
                if (!a->busy && a->ready) {
                    pick = a;
                    rr = (rr + q + 1) % active.size();
                    break;
                }
            }
            if (pick) {
                pick->busy = true;
                const ovs_job &job = jobs[pick->job];
                const int pts = pick->n_points;
                lk.unlock();
                const Clock::time_point s0 = Clock::now();
                const int r = run_step(job, *pick);
                const double ms = since(s0) * 1e3;
                lk.lock();
                step_ms.push_back(ms);
                if (r) {
                    if (res) res[pick->job].failed = r;
                    close_job(lk, pick);
                    continue;
                }
                pick->steps_done++;
                point_steps += pts;
                if (res) res[pick->job].point_steps += pts;
                if (pick->steps_done >= job.steps) {
                    close_job(lk, pick);
                } else {
                    pick->busy = false;
                    cv.notify_all();
                }
                continue;
            }
            // nothing to do now: wait for the next arrival or a released job
            if (next < n_jobs && !arrived(next) && active.empty())
                cv.wait_until(lk, t0 + std::chrono::duration_cast<Clock::duration>(
                                           std::chrono::duration<double>(jobs[next].arrival_s)));
            else if (next < n_jobs && !arrived(next))
                cv.wait_for(lk, std::chrono::milliseconds(1));
            else
                cv.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
};

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const double pos = q / 100.0 * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" double ovs_job_bytes(int n_slots)
{
    // per slot: state and RK4/Biot-Savart scratch (~300 B), tree with its
    // sort buffers (~260 B per point, sized by slots), topology scratch
    // (~60 B); plus a fixed 16 MiB of context buffers
    return 620.0 * (double)n_slots + 16.0 * 1024 * 1024;
}

extern "C" int ovs_run(const ovs_job *jobs, int n_jobs, int workers, int max_active,
                       double hbm_cap_bytes, int realtime, ovs_report *report,
                       ovs_result *results)
{
    if (n_jobs < 0 || (n_jobs > 0 && !jobs) || workers < 1 || max_active < 1 ||
        !(hbm_cap_bytes > 0) || !report)
        return OVG_EINVAL;
    for (int j = 1; j < n_jobs; ++j)
        if (jobs[j].arrival_s < jobs[j - 1].arrival_s) return OVG_EINVAL;
    if (results) std::memset(results, 0, sizeof(ovs_result) * (size_t)n_jobs);
    Server S;
    S.jobs = jobs;
    S.n_jobs = n_jobs;
    S.max_active = max_active;
    S.cap = hbm_cap_bytes;
    S.realtime = realtime != 0;
    S.res = results;
    S.t0 = Clock::now();
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w) pool.emplace_back([&S] { S.worker(); });
    for (auto &t : pool) t.join();
    std::memset(report, 0, sizeof *report);
    report->wall_s = since(S.t0);
    report->point_steps = S.point_steps;
    report->point_steps_per_s = report->wall_s > 0 ? S.point_steps / report->wall_s : 0;
    report->step_p50_ms = percentile(S.step_ms, 50);
    report->step_p99_ms = percentile(S.step_ms, 99);
    report->steps_total = (int)S.step_ms.size();
    report->peak_bytes = S.peak_bytes;
    report->peak_active = S.peak_active;
    std::vector<double> lat;
    for (int j = 0; j < n_jobs && results; ++j) {
        if (results[j].failed) {
            report->jobs_failed++;
        } else {
            report->jobs_done++;
            lat.push_back(results[j].latency_s);
        }
    }
    report->job_p50_s = percentile(lat, 50);
    report->job_p99_s = percentile(lat, 99);
    return OVG_OK;
}
