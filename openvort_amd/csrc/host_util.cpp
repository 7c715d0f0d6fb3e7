// Host-side utilities of the C ABI: error text, parameter defaults, the
// periodic image tables and the seeded ring-tangle generator.  No GPU code.
#include "../../include/openvort_gpu.h"
#include "engine_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

namespace ovg {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error()
{
    g_last_error.clear();
    (void)hipGetLastError();
}

int launch_status(int r)
{
    if (r != OVG_OK) return r;
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(OVG_EDEVICE, "kernel launch failed: %s", hipGetErrorString(e));
    return OVG_OK;
}

}  // namespace ovg

extern "C" const char *ovg_last_error(void) { return ovg::g_last_error.c_str(); }

extern "C" const char *ovg_version(void) { return "openvort-mi355x 0.1 (gfx950)"; }

extern "C" size_t ovg_params_sizeof(void) { return sizeof(ovg_params); }
extern "C" size_t ovg_tune_sizeof(void) { return sizeof(ovg_tune); }

extern "C" size_t ovg_stats_sizeof(void) { return sizeof(ovg_stats); }

// Periodic image tables of src/boundary_images.c:13-72: every integer shift
// in {-1,0,1}^3 except zero with at most 1 (6 images), 2 (18) or 3 (26)
// nonzero components, in lexicographic order of (sx, sy, sz) -- the order
// of the reference tables.
extern "C" int ovg_params_set_periodic_images(ovg_params *p, int n)
{
    if (!p) return ovg::set_error(OVG_EINVAL, "params is NULL");
    int maxnz;
    switch (n) {
    case 0: p->n_images = 0; return OVG_OK;
    case 6: maxnz = 1; break;
    case 18: maxnz = 2; break;
    case 26: maxnz = 3; break;
    default: return ovg::set_error(OVG_EINVAL, "periodic image count must be 0, 6, 18 or 26, got %d", n);
    }
    int j = 0;
    for (int sx = -1; sx <= 1; ++sx)
        for (int sy = -1; sy <= 1; ++sy)
            for (int sz = -1; sz <= 1; ++sz) {
                int nz = (sx != 0) + (sy != 0) + (sz != 0);
                if (nz == 0 || nz > maxnz) continue;
                p->images[j][0] = sx;
                p->images[j][1] = sy;
                p->images[j][2] = sz;
                ++j;
            }
    p->n_images = j;
    for (int i = 0; i < 26; ++i) p->image_reflect[i] = 0;
    return OVG_OK;
}

// Mirror-wall image tables of src/boundary_images.c:75-230, entry by entry
// (the order is the order of the image sums): shift (sx, sy, sz) and the
// mask of the faces mirrored through.  Kept as the reference has them,
// quirks included: wall_1_18's {1,-1,0} entry has number_of_reflects 1 with
// an empty (zero-filled) list, i.e. LEFT; wall_parallel_planes lists
// {-1,1,-1} and {-1,1,1} twice and {1,-1,-1}, {1,-1,1} not at all.
namespace {
constexpr int L_ = 1 << OVG_FACE_LEFT, B_ = 1 << OVG_FACE_BACK, F_ = 1 << OVG_FACE_FRONT,
              D_ = 1 << OVG_FACE_DOWN, U_ = 1 << OVG_FACE_UP;
struct WallImage {
    signed char s[3];
    int mask;
};
const WallImage kWall1Open[] = {{{0, 0, -1}, D_}};
const WallImage kWall16[] = {{{-1, 0, 0}, 0}, {{0, -1, 0}, 0}, {{0, 0, -1}, D_},
                             {{0, 1, 0}, 0},  {{1, 0, 0}, 0}};
const WallImage kWall118[] = {
    {{-1, -1, 0}, 0}, {{-1, 0, -1}, D_}, {{-1, 0, 0}, 0},  {{-1, 0, 1}, 0},  {{-1, 1, 0}, 0},
    {{0, -1, -1}, D_}, {{0, -1, 0}, 0},  {{0, -1, 1}, 0},  {{0, 0, -1}, D_}, {{0, 0, 1}, 0},
    {{0, 1, -1}, D_},  {{0, 1, 0}, 0},   {{0, 1, 1}, 0},   {{1, -1, 0}, L_}, {{1, 0, -1}, D_},
    {{1, 0, 0}, 0},    {{1, 0, 1}, 0},   {{1, 1, 0}, 0}};
const WallImage kWall126[] = {
    {{-1, -1, -1}, D_}, {{-1, -1, 0}, 0}, {{-1, -1, 1}, 0}, {{-1, 0, -1}, D_}, {{-1, 0, 0}, 0},
    {{-1, 0, 1}, 0},    {{-1, 1, -1}, D_}, {{-1, 1, 0}, 0}, {{-1, 1, 1}, 0},   {{0, -1, -1}, D_},
    {{0, -1, 0}, 0},    {{0, -1, 1}, 0},  {{0, 0, -1}, D_}, {{0, 0, 1}, 0},    {{0, 1, -1}, D_},
    {{0, 1, 0}, 0},     {{0, 1, 1}, 0},   {{1, -1, -1}, D_}, {{1, -1, 0}, 0},  {{1, -1, 1}, 0},
    {{1, 0, -1}, D_},   {{1, 0, 0}, 0},   {{1, 0, 1}, 0},   {{1, 1, -1}, D_},  {{1, 1, 0}, 0},
    {{1, 1, 1}, 0}};
const WallImage kWall24[] = {{{-1, 0, 0}, 0}, {{1, 0, 0}, 0},   {{0, -1, 0}, 0},
                             {{0, 1, 0}, 0},  {{0, 0, -1}, D_}, {{0, 0, 1}, U_}};
const WallImage kWall22[] = {{{-1, 0, 0}, 0}, {{1, 0, 0}, 0},   {{0, -1, 0}, B_},
                             {{0, 1, 0}, F_}, {{0, 0, -1}, D_}, {{0, 0, 1}, U_}};
const WallImage kWallPlanes[] = {
    {{-1, 0, 0}, 0},  {{1, 0, 0}, 0},   {{0, -1, 0}, 0},  {{0, 1, 0}, 0},   {{0, 0, -1}, D_},
    {{0, 0, 1}, U_},  {{-1, -1, 0}, 0}, {{1, -1, 0}, 0},  {{-1, 1, 0}, 0},  {{1, 1, 0}, 0},
    {{1, 0, -1}, D_}, {{-1, 0, -1}, D_}, {{0, 1, -1}, D_}, {{0, -1, -1}, D_}, {{1, 0, 1}, U_},
    {{-1, 0, 1}, U_}, {{0, 1, 1}, U_},  {{0, -1, 1}, U_}, {{1, 1, -1}, D_}, {{-1, 1, -1}, D_},
    {{-1, 1, -1}, D_}, {{-1, -1, -1}, D_}, {{1, 1, 1}, U_}, {{-1, 1, 1}, U_}, {{-1, 1, 1}, U_},
    {{-1, -1, 1}, U_}};
const WallImage kWallCanal[] = {
    {{-1, 0, 0}, 0},       {{1, 0, 0}, 0},        {{0, -1, 0}, B_},      {{0, 1, 0}, F_},
    {{0, 0, -1}, D_},      {{0, 0, 1}, U_},       {{0, -1, -1}, B_ | D_}, {{0, -1, 1}, B_ | U_},
    {{0, 1, -1}, F_ | D_}, {{0, 1, 1}, F_ | U_},  {{-1, -1, 0}, B_},     {{-1, 1, 0}, F_},
    {{-1, 0, -1}, D_},     {{-1, 0, 1}, U_},      {{-1, -1, -1}, B_ | D_}, {{-1, -1, 1}, B_ | U_},
    {{-1, 1, -1}, F_ | D_}, {{-1, 1, 1}, F_ | U_}, {{1, -1, 0}, B_},      {{1, 1, 0}, F_},
    {{1, 0, -1}, D_},      {{1, 0, 1}, U_},       {{1, -1, -1}, B_ | D_}, {{1, -1, 1}, B_ | U_},
    {{1, 1, -1}, F_ | D_}, {{1, 1, 1}, F_ | U_}};
}  // namespace

extern "C" int ovg_params_set_wall_images(ovg_params *p, int config)
{
    if (!p) return ovg::set_error(OVG_EINVAL, "params is NULL");
    const WallImage *t = nullptr;
    int n = 0;
    switch (config) {
#define OVG_TABLE_(id, arr) case id: t = arr; n = (int)(sizeof(arr) / sizeof(arr[0])); break;
    OVG_TABLE_(OVG_IMAGES_WALL_1_OPEN, kWall1Open)
    OVG_TABLE_(OVG_IMAGES_WALL_1_6, kWall16)
    OVG_TABLE_(OVG_IMAGES_WALL_1_18, kWall118)
    OVG_TABLE_(OVG_IMAGES_WALL_1_26, kWall126)
    OVG_TABLE_(OVG_IMAGES_WALL_2_4, kWall24)
    OVG_TABLE_(OVG_IMAGES_WALL_2_2, kWall22)
    OVG_TABLE_(OVG_IMAGES_WALL_PARALLEL_PLANES, kWallPlanes)
    OVG_TABLE_(OVG_IMAGES_WALL_CANAL, kWallCanal)
#undef OVG_TABLE_
    default: return ovg::set_error(OVG_EINVAL, "unknown wall image table %d", config);
    }
    for (int j = 0; j < 26; ++j) {
        for (int k = 0; k < 3; ++k) p->images[j][k] = j < n ? t[j].s[k] : 0;
        p->image_reflect[j] = j < n ? t[j].mask : 0;
    }
    p->n_images = n;
    return OVG_OK;
}

// Defaults of src/initial_conditions.c:12-91, periodic cube [0,D]^3.
extern "C" void ovg_params_default(ovg_params *p, double D)
{
    std::memset(p, 0, sizeof *p);
    p->kappa = 9.97e-4;
    p->vortex_width = 1e-8;
    p->use_mutual_friction = 1;
    p->alpha = 0.034;
    p->alpha_p = 1.383e-2;
    p->vn.type = OVG_FLOW_NONE;
    p->vs.type = OVG_FLOW_NONE;
    p->vb.type = OVG_FLOW_NONE;
    for (int k = 0; k < 3; ++k) {
        p->box_lo[k] = 0.0;
        p->box_hi[k] = D;
    }
    for (int k = 0; k < 6; ++k) p->wall[k] = OVG_WALL_PERIODIC;
    ovg_params_set_periodic_images(p, 26);
    p->method = OVG_METHOD_BH;
    p->precision = OVG_PREC_FP64;
    p->theta = 0.4;
    p->leaf_size = 1;
    p->traverse = OVG_TRAVERSE_POINT;
    p->rec_dist = 4e-4;
    p->rec_angle = 5 * M_PI / 180.0;
    p->freq_to_cutoff = 4.0 / 5;
    p->rk4_classic = 0;
}

// ---- seeded generator ---------------------------------------------------
namespace {

struct Rng {
    uint64_t s;
    uint64_t next()
    {  // SplitMix64 (Steele, Lea, Flood 2014)
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
    void normals(double *z, int n)
    {  // Box-Muller pairs
        for (int i = 0; i < n; i += 2) {
            double u1 = 1.0 - u01(), u2 = u01();
            double r = std::sqrt(-2.0 * std::log(u1));
            z[i] = r * std::cos(2.0 * M_PI * u2);
            if (i + 1 < n) z[i + 1] = r * std::sin(2.0 * M_PI * u2);
        }
    }
};

struct Ring {
    double R, c[3], u[3], v[3];
    int n;
};

Ring make_ring(uint64_t seed, int k, double D, double r_min, double r_max, double dxi)
{
    Rng g{seed ^ (0x632BE59BD9B4E019ULL * (uint64_t)(k + 1))};
    g.next();
    Ring r;
    r.R = r_min + (r_max - r_min) * g.u01();
    for (int i = 0; i < 3; ++i) r.c[i] = D * g.u01();
    double nv[4], len = 0;
    do {
        g.normals(nv, 4);
        len = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    } while (len < 1e-12);
    double nrm[3] = {nv[0] / len, nv[1] / len, nv[2] / len};
    // in-plane basis: u = a - (a.n) n normalised, v = n x u
    double a[3] = {1, 0, 0};
    if (std::fabs(nrm[0]) > 0.9) { a[0] = 0; a[1] = 1; }
    double an = a[0] * nrm[0] + a[1] * nrm[1] + a[2] * nrm[2];
    double u[3] = {a[0] - an * nrm[0], a[1] - an * nrm[1], a[2] - an * nrm[2]};
    double ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int i = 0; i < 3; ++i) r.u[i] = u[i] / ul;
    r.v[0] = nrm[1] * r.u[2] - nrm[2] * r.u[1];
    r.v[1] = nrm[2] * r.u[0] - nrm[0] * r.u[2];
    r.v[2] = nrm[0] * r.u[1] - nrm[1] * r.u[0];
    r.n = (int)std::floor(2.0 * M_PI * r.R / dxi);
    if (r.n < 3) r.n = 3;
    return r;
}

double wrap01(double x, double D)
{
    double w = std::fmod(x, D);
    if (w < 0) w += D;
    if (w >= D) w -= D;
    return w;
}

}  // namespace

extern "C" int ovg_generate_rings(uint64_t seed, int n_rings, double D, double r_min,
                                  double r_max, double dxi, int n_cap, double *vnodes,
                                  int *connections, int *status)
{
    if (n_rings < 0 || !(D > 0) || !(dxi > 0) || r_min > r_max || n_cap < 0) {
        ovg::set_error(OVG_EINVAL, "ovg_generate_rings: bad arguments");
        return -1;
    }
    long total = 0;
    for (int k = 0; k < n_rings; ++k) {
        Ring r = make_ring(seed, k, D, r_min, r_max, dxi);
        if (total + r.n > n_cap) {
            if (vnodes || connections || status) {
                ovg::set_error(OVG_ELIMIT, "ovg_generate_rings: %ld points exceed capacity %d",
                               total + r.n, n_cap);
                return -1;
            }
        } else if (vnodes || connections || status) {
            for (int j = 0; j < r.n; ++j) {
                long s = total + j;
                double phi = 2.0 * M_PI * j / r.n;
                double cph = std::cos(phi), sph = std::sin(phi);
                if (vnodes)
                    for (int i = 0; i < 3; ++i)
                        vnodes[3 * s + i] = wrap01(r.c[i] + r.R * (cph * r.u[i] + sph * r.v[i]), D);
                if (connections) {
                    connections[2 * s] = (int)(total + (j + 1) % r.n);
                    connections[2 * s + 1] = (int)(total + (j + r.n - 1) % r.n);
                }
                if (status) {
                    status[2 * s] = OVG_FREE;
                    status[2 * s + 1] = -1;
                }
            }
        }
        total += r.n;
    }
    if (vnodes || connections || status) {
        for (long s = total; s < n_cap; ++s) {
            if (vnodes) vnodes[3 * s] = vnodes[3 * s + 1] = vnodes[3 * s + 2] = 0.0;
            if (connections) connections[2 * s] = connections[2 * s + 1] = -1;
            if (status) {
                status[2 * s] = OVG_EMPTY;
                status[2 * s + 1] = -1;
            }
        }
    }
    return total > 0x7fffffff ? -1 : (int)total;
}

extern "C" int ovg_rings_for_points(uint64_t seed, int target_points, double D, double r_min,
                                    double r_max, double dxi)
{
    long total = 0;
    int k = 0;
    while (total < target_points) {
        total += make_ring(seed, k, D, r_min, r_max, dxi).n;
        ++k;
        if (k > 100000000) return -1;
    }
    return k;
}
