// Vortex injection (src/vortex_injection.c): the geometry of add_loop
// (:44-103, with perpendicular :11-35) and the random schedules of
// inject_loop (:711-745) and inject_line_pairs (:760-800), host-only.
// ovg_add_loop and ovg_add_line_z (engine.hip) place the points of a loop
// or of a line pinned to the DOWN and UP walls in a device-resident tangle.
//
// The reference draws from the process-wide drand48 stream and keeps the
// last injection time in a static; here both are the caller's: the 48-bit
// state in erand48 form (srand48(s) is {0x330E, s & 0xffff, (s >> 16) &
// 0xffff}) and last_injection.  Vector arithmetic follows src/vec3_math.c
// operation by operation, so the points are bit-identical to the reference's.
#include <cmath>

#include "openvort_gpu.h"

namespace {

struct V3 {
    double p[3];
};

// vec3_dot, vec3_len (:91-96, :167-174): sums from 0 in index order
double dot(const V3 &u, const V3 &v)
{
    double out = 0;
    for (int k = 0; k < 3; ++k) out += u.p[k] * v.p[k];
    return out;
}
double len(const V3 &u)
{
    double out = 0;
    for (int k = 0; k < 3; ++k) out += u.p[k] * u.p[k];
    return std::sqrt(out);
}
V3 mul(const V3 &u, double m)
{
    V3 r;
    for (int k = 0; k < 3; ++k) r.p[k] = u.p[k] * m;
    return r;
}
V3 add(const V3 &u, const V3 &v)
{
    V3 r;
    for (int k = 0; k < 3; ++k) r.p[k] = u.p[k] + v.p[k];
    return r;
}
V3 cross(const V3 &u, const V3 &v)  // vec3_cross :119-127
{
    V3 r;
    r.p[0] = u.p[1] * v.p[2] - u.p[2] * v.p[1];
    r.p[1] = u.p[2] * v.p[0] - u.p[0] * v.p[2];
    r.p[2] = u.p[0] * v.p[1] - u.p[1] * v.p[0];
    return r;
}
// vec3_normalize :208-217: a zero vector is left as it is
void normalize(V3 &v)
{
    if (!(v.p[0] == 0 && v.p[1] == 0 && v.p[2] == 0)) {
        double d = len(v);
        v = mul(v, 1 / d);
    }
}
// perpendicular, src/vortex_injection.c:11-35
V3 perpendicular(const V3 &dir)
{
    const V3 a = {{1, 0, 0}}, b = {{0, 1, 0}};
    V3 res = add(mul(dir, -dot(a, dir)), a);
    if (len(res) > 1e-8) {
        normalize(res);
        return res;
    }
    res = add(mul(dir, -dot(b, dir)), b);
    normalize(res);
    return res;
}

// erand48: X <- (0x5DEECE66D X + 0xB) mod 2^48, returned as X / 2^48
double rand48(unsigned short s[3])
{
    unsigned long long x = (unsigned long long)s[0] | ((unsigned long long)s[1] << 16) |
                           ((unsigned long long)s[2] << 32);
    x = (0x5DEECE66DULL * x + 0xBULL) & ((1ULL << 48) - 1);
    s[0] = (unsigned short)x;
    s[1] = (unsigned short)(x >> 16);
    s[2] = (unsigned short)(x >> 32);
    return std::ldexp((double)x, -48);
}

}  // namespace

extern "C" int ovg_loop_points(const double center[3], const double dir[3], double r,
                               double dl_min, double dl_max, double *xyz, int cap, int *n_points)
{
    if (!center || !dir || !n_points || (cap > 0 && !xyz) || !(dl_min + dl_max > 0))
        return OVG_EINVAL;
    const int npts = (int)(4 * M_PI * r / (dl_max + dl_min));
    *n_points = npts;
    V3 c = {{center[0], center[1], center[2]}}, d = {{dir[0], dir[1], dir[2]}};
    // the circle's base (:62-67)
    const V3 zdir = perpendicular(d);
    V3 u = cross(zdir, d), v = cross(d, u);
    normalize(u);
    normalize(v);
    for (int k = 0; k < npts && k < cap; ++k) {  // :73-86
        const double phi = 2 * M_PI / npts * k;
        // cos(phi) and sin(phi) as the reference build computes them: gcc
        // fuses the two calls into glibc's sincos, whose sine can differ
        // from sin()'s by an ulp
        double sn, cs;
        ::sincos(phi, &sn, &cs);
        V3 p = add(add(mul(u, r * cs), mul(v, r * sn)), c);
        for (int q = 0; q < 3; ++q) xyz[3 * (size_t)k + q] = p.p[q];
    }
    return OVG_OK;
}

extern "C" int ovg_inject_loop_draw(double t, double frequency, const double box_lo[3],
                                    const double box_hi[3], unsigned short rng[3],
                                    double *last_injection, double center[3], double dir[3],
                                    double *r, int *due)
{
    if (!box_lo || !box_hi || !rng || !last_injection || !center || !dir || !r || !due)
        return OVG_EINVAL;
    // vec3(2*(drand48() - 0.5), 2*(drand48() - 0.5), -drand48()) (:718): the
    // reference build evaluates the three arguments right to left
    const double d3 = rand48(rng), d2 = rand48(rng), d1 = rand48(rng);
    V3 dv = {{2 * (d1 - 0.5), 2 * (d2 - 0.5), -d3}};
    normalize(dv);
    *due = 0;
    if (t - *last_injection < 1 / frequency) return OVG_OK;  // :723
    *last_injection = t;
    *due = 1;
    const double left = box_lo[0], right = box_hi[0], back = box_lo[1], front = box_hi[1];
    center[2] = box_hi[2];  // :735-737
    center[0] = left + (right - left) * rand48(rng);
    center[1] = back + (front - back) * rand48(rng);
    const double D = std::fabs(right - left);  // :740-744
    const double rmin = 0.05 * D, rmax = 0.25 * D;
    *r = rmin + (rmax - rmin) * rand48(rng);
    for (int q = 0; q < 3; ++q) dir[q] = dv.p[q];
    return OVG_OK;
}

extern "C" int ovg_inject_line_pairs_draw(double t, double frequency, int n_pairs, int polarized,
                                          const double box_lo[3], const double box_hi[3],
                                          unsigned short rng[3], double *last_injection,
                                          double *xy, int *due)
{
    if (!box_lo || !box_hi || !rng || !last_injection || !due || n_pairs < 0 ||
        (n_pairs > 0 && !xy))
        return OVG_EINVAL;
    *due = 0;
    if (t - *last_injection < 1 / frequency) return OVG_OK;  // :772-773
    *last_injection = t;
    *due = 1;
    const double left = box_lo[0], right = box_hi[0], back = box_lo[1], front = box_hi[1];
    for (int k = 0; k < n_pairs; ++k) {  // :776-799
        const double x1 = left + rand48(rng) * (right - left);
        const double x2 = left + rand48(rng) * (right - left);
        double y1, y2;
        if (polarized) {
            y1 = back + (1 + rand48(rng)) * (front - back) / 2;  // positive line in (0.5, 1)
            y2 = back + rand48(rng) * (front - back) / 2;        // negative line in (0, 0.5)
        } else {
            y1 = back + rand48(rng) * (front - back);
            y2 = back + rand48(rng) * (front - back);
        }
        xy[4 * (size_t)k] = x1;
        xy[4 * (size_t)k + 1] = y1;
        xy[4 * (size_t)k + 2] = x2;
        xy[4 * (size_t)k + 3] = y2;
    }
    return OVG_OK;
}
