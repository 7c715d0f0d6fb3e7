// Tangle files of the reference (src/utils.c:68-199): one text line per
// point, 18 tab-separated columns
//   loop  x y z  vels(3)  tangents(3)  normals(3)  slot reverse forward status pin_wall
// written loop by loop in the slot order of each loop's first point.
// Host code; the context entry points download / upload through the C ABI.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/openvort_gpu.h"
#include "engine_internal.h"

namespace {

struct Arrays {
    int n;
    const double *x, *v, *t, *nrm;
    const int *conn, *st;
};

void put_vec(FILE *f, const double *p, int digits)
{
    std::fprintf(f, "%.*g\t%.*g\t%.*g", digits, p[0], digits, p[1], digits, p[2]);
}

// save_point, src/utils.c:74-91
void put_point(FILE *f, int loop, const Arrays &a, int i, int digits)
{
    static const double zero[3] = {0, 0, 0};
    std::fprintf(f, "%d\t", loop);
    put_vec(f, a.x + 3 * (size_t)i, digits);
    std::fprintf(f, "\t");
    put_vec(f, a.v ? a.v + 3 * (size_t)i : zero, digits);
    std::fprintf(f, "\t");
    put_vec(f, a.t ? a.t + 3 * (size_t)i : zero, digits);
    std::fprintf(f, "\t");
    put_vec(f, a.nrm ? a.nrm + 3 * (size_t)i : zero, digits);
    std::fprintf(f, "\t%d\t%d\t%d\t%d\t%d\n", i, a.conn[2 * (size_t)i + 1], a.conn[2 * (size_t)i],
                 a.st[2 * (size_t)i], a.st[2 * (size_t)i + 1]);
}

// save_tangle, src/utils.c:132-184, including its traversal: forward from
// the first unvisited live slot until the link returns to it; an open end
// turns back along the reverse links; the last point is written only when
// its slot is > 0.
int write_file(const char *path, const Arrays &a, int digits)
{
    FILE *f = std::fopen(path, "w");
    if (!f) return ovg::set_error(OVG_EINVAL, "cannot open %s for writing", path);
    std::vector<char> visited((size_t)a.n, 0);
    int loop = 0;
    auto fwd = [&](int k) { return a.conn[2 * (size_t)k]; };
    auto rev = [&](int k) { return a.conn[2 * (size_t)k + 1]; };
    for (int k = 0; k < a.n; ++k) {
        if (visited[k]) continue;
        if (a.st[2 * (size_t)k] == OVG_EMPTY) {
            visited[k] = 1;
            continue;
        }
        const int first = k;
        int curr = k;
        long guard = 0;
        while (fwd(curr) != first) {
            put_point(f, loop, a, curr, digits);
            visited[curr] = 1;
            curr = fwd(curr);
            if (curr < 0) {
                curr = rev(first);
                while (curr > 0 && rev(curr) > 0) {
                    put_point(f, loop, a, curr, digits);
                    visited[curr] = 1;
                    curr = rev(curr);
                }
                break;
            }
            if (++guard > a.n) {
                std::fclose(f);
                return ovg::set_error(OVG_EINVAL, "links of slot %d do not close", first);
            }
        }
        if (curr > 0) {
            put_point(f, loop, a, curr, digits);
            visited[curr] = 1;
        }
        loop++;
    }
    std::fclose(f);
    return OVG_OK;
}

struct Point {
    double x[3], v[3], t[3], n[3];
    int id, reverse, forward, status, pin_wall;
};

// load_point, src/utils.c:93-130: the same fscanf conversion
int read_points(const char *path, std::vector<Point> &pts)
{
    FILE *f = std::fopen(path, "r");
    if (!f) return ovg::set_error(OVG_EINVAL, "cannot open %s", path);
    for (;;) {
        Point p;
        int loop;
        int r = std::fscanf(
            f, "%d\t%lg\t%lg\t%lg\t%lg\t%lg\t%lg\t%lg\t%lg\t%lg\t%lg\t%lg\t%lg\t%d\t%d\t%d\t%d\t%d",
            &loop, &p.x[0], &p.x[1], &p.x[2], &p.v[0], &p.v[1], &p.v[2], &p.t[0], &p.t[1],
            &p.t[2], &p.n[0], &p.n[1], &p.n[2], &p.id, &p.reverse, &p.forward, &p.status,
            &p.pin_wall);
        if (r == EOF) break;
        if (r != 18 || p.id < 0) {
            std::fclose(f);
            return ovg::set_error(OVG_EINVAL, "%s: malformed line after %zu points", path,
                                  pts.size());
        }
        pts.push_back(p);
    }
    std::fclose(f);
    return OVG_OK;
}

// slots after load_tangle on a tangle created with create_tangle(512)
// (src/main.c:85): doubled while an id does not fit (src/utils.c:117-119)
int slots_for(const std::vector<Point> &pts)
{
    int n = 512;
    for (const Point &p : pts)
        while (p.id > n - 1) n *= 2;
    return n;
}

}  // namespace

extern "C" int ovg_write_tangle_file(const char *path, int n, const double *vnodes,
                                     const double *vels, const double *tangents,
                                     const double *normals, const int *connections,
                                     const int *status, int digits)
{
    ovg::clear_error();
    if (!path || n < 0 || (n > 0 && (!vnodes || !connections || !status)) || digits < 1 ||
        digits > 17)
        return ovg::set_error(OVG_EINVAL, "bad arguments to ovg_write_tangle_file");
    Arrays a{n, vnodes, vels, tangents, normals, connections, status};
    return write_file(path, a, digits);
}

extern "C" int ovg_read_tangle_file(const char *path, int n_cap, int *n_slots, double *vnodes,
                                    double *vels, double *tangents, double *normals,
                                    int *connections, int *status)
{
    ovg::clear_error();
    if (!path || !n_slots) return ovg::set_error(OVG_EINVAL, "bad arguments");
    std::vector<Point> pts;
    OVG_TRY(read_points(path, pts));
    const int n = slots_for(pts);
    *n_slots = n;
    if (!vnodes && !vels && !tangents && !normals && !connections && !status) return OVG_OK;
    if (n_cap < n) return ovg::set_error(OVG_ELIMIT, "the file needs %d slots, %d given", n, n_cap);
    for (int k = 0; k < n; ++k) {
        for (int q = 0; q < 3; ++q) {
            if (vnodes) vnodes[3 * (size_t)k + q] = 0;
            if (vels) vels[3 * (size_t)k + q] = 0;
            if (tangents) tangents[3 * (size_t)k + q] = 0;
            if (normals) normals[3 * (size_t)k + q] = 0;
        }
        if (connections) connections[2 * (size_t)k] = connections[2 * (size_t)k + 1] = -1;
        if (status) {
            status[2 * (size_t)k] = OVG_EMPTY;
            status[2 * (size_t)k + 1] = -1;  // NOT_A_FACE
        }
    }
    for (const Point &p : pts) {
        const size_t k = (size_t)p.id;
        for (int q = 0; q < 3; ++q) {
            if (vnodes) vnodes[3 * k + q] = p.x[q];
            if (vels) vels[3 * k + q] = p.v[q];
            if (tangents) tangents[3 * k + q] = p.t[q];
            if (normals) normals[3 * k + q] = p.n[q];
        }
        if (connections) {
            connections[2 * k] = p.forward;
            connections[2 * k + 1] = p.reverse;
        }
        if (status) {
            status[2 * k] = p.status;
            status[2 * k + 1] = p.pin_wall;
        }
    }
    return OVG_OK;
}

extern "C" int ovg_save_tangle(ovg_ctx *ctx, const char *path, int digits)
{
    ovg::clear_error();
    if (!ctx || !path) return ovg::set_error(OVG_EINVAL, "bad arguments to ovg_save_tangle");
    if (ovg_forest_size(ctx) > 1)
        return ovg::set_error(OVG_EUNSUPPORTED, "ovg_save_tangle: the context holds a forest; "
                              "a tangle file holds one tangle");
    const int n = ovg_num_slots(ctx);
    if (n < 0) return ovg::set_error(OVG_EINVAL, "no tangle");
    std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n), t(3 * (size_t)n), nr(3 * (size_t)n);
    std::vector<int> conn(2 * (size_t)n), st(2 * (size_t)n);
    OVG_TRY(ovg_download(ctx, x.data(), v.data(), nullptr, t.data(), nr.data()));
    OVG_TRY(ovg_download_connections(ctx, conn.data(), st.data()));
    return ovg_write_tangle_file(path, n, x.data(), v.data(), t.data(), nr.data(), conn.data(),
                                 st.data(), digits);
}

extern "C" int ovg_load_tangle(ovg_ctx *ctx, const char *path)
{
    ovg::clear_error();
    if (!ctx || !path) return ovg::set_error(OVG_EINVAL, "bad arguments to ovg_load_tangle");
    int n = 0;
    OVG_TRY(ovg_read_tangle_file(path, 0, &n, nullptr, nullptr, nullptr, nullptr, nullptr,
                                 nullptr));
    std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n), t(3 * (size_t)n), nr(3 * (size_t)n);
    std::vector<int> conn(2 * (size_t)n), st(2 * (size_t)n);
    OVG_TRY(ovg_read_tangle_file(path, n, &n, x.data(), v.data(), t.data(), nr.data(),
                                 conn.data(), st.data()));
    OVG_TRY(ovg_upload(ctx, n, x.data(), conn.data(), st.data()));
    OVG_TRY(ovg_upload_field(ctx, OVG_FIELD_VELS, v.data()));
    OVG_TRY(ovg_upload_field(ctx, OVG_FIELD_TANGENTS, t.data()));
    OVG_TRY(ovg_upload_field(ctx, OVG_FIELD_NORMALS, nr.data()));
    return OVG_OK;
}
