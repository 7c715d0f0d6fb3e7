// Device-side primitives shared by the kernels.  Everything that decides a
// reference guard (seg_pwrap, the segment-field formula and its < 1e-8
// tests) is written in the reference's operation order; the library is
// compiled with -ffp-contract=off so nothing is fused behind our back.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ovg {

constexpr int kMaxImages = 26;
constexpr int kMaxLevel = 21;  // octree levels encoded in a 63-bit path key

// Per-tangle constants, precomputed on the host exactly as the reference
// computes them inside its loops.
struct TangleParams {
    double lo[3], hi[3];
    double L[3];        // hi - lo                       (src/boundaries.c:61-63)
    double halfL[3];    // L/2                            (src/boundaries.c:68-79)
    int periodic[3];    // wall[LEFT], wall[BACK], wall[DOWN] are periodic
    int wrap_face[6];   // wall[face] == periodic, for enforce_periodic
    int n_images;
    int pad0;
    double Ls[kMaxImages][3];  // L * shift, the image offsets (src/tangle.c:378)
    double W2[kMaxImages][3];  // 2 * wall coordinate on mirrored axes (src/tangle.c:385-400)
    int refl[kMaxImages];      // per image: bit a set = mirrored on axis a
    double f_bs;        // KAPPA / 4 / M_PI               (src/tangle.c:545)
    double kappa, a0, sqrtE;
    double evs[3];      // external superfluid flow at this call's t
    double evn[3];      // external normal flow
    double evb[3];      // boundary flow (PINNED nodes)
    double alpha, alpha_p;
    int use_mf;
    int pad1;
    double theta;
    double crit_scale;  // 1/theta^2 for the squared acceptance test
    int leaf_size;
    int group_level;    // list traversal: octree level that bounds a target group
    double near2;       // mixed precision: FP64 zone, squared (cm^2)
    // tuned trees: fixed-point units 2^-shift cm for this tangle's size
    double fix_scale, fix_inv, fix_lim;
    // Forests (many tangles in one context): entry 0 of the parameter table
    // holds the tangle count and the per-slot tangle index; every entry
    // holds its own tangle's constants.  n_tangles == 1: one tangle.
    int n_tangles;
    int pad2;
    const int *tid;
};

// Constants of the tangle that slot k belongs to.
__device__ __forceinline__ const TangleParams &tparams_of(const TangleParams *Pp, int k)
{
    return Pp[0].n_tangles > 1 ? Pp[Pp[0].tid[k]] : Pp[0];
}

// Tuned-tree keys (list and group traversals): the tangle index above 48
// Morton bits at key bits 1..48 (16 levels of 3-bit digits, level 1 most
// significant); bit 0 is zero.
constexpr int kKeyTangleShift = 49;
__device__ __host__ __forceinline__ int key_tangle(uint64_t key)
{
    return (int)(key >> kKeyTangleShift);
}

struct Vec3 {
    double x, y, z;
};

// shifted, src/tangle.c:370-414: r - L * shift, then 2 * wall - r (the
// unshifted r) on each mirrored axis.
__host__ __device__ __forceinline__ void image_point(const TangleParams &P, int im, double rx,
                                                     double ry, double rz, double &qx,
                                                     double &qy, double &qz)
{
    qx = rx - P.Ls[im][0];
    qy = ry - P.Ls[im][1];
    qz = rz - P.Ls[im][2];
    const int m = P.refl[im];
    if (m) {
        if (m & 1) qx = P.W2[im][0] - rx;
        if (m & 2) qy = P.W2[im][1] - ry;
        if (m & 4) qz = P.W2[im][2] - rz;
    }
}

// mirror_vector_reflect, src/boundaries.c:201-211, for each mirrored axis:
// v + n (-2 v.n) with n = +-e_a changes only component a, to v_a + (-2 v_a).
__host__ __device__ __forceinline__ void image_reflect(const TangleParams &P, int im, double &vx,
                                                       double &vy, double &vz)
{
    const int m = P.refl[im];
    if (m) {
        if (m & 1) vx = vx + -2 * vx;
        if (m & 2) vy = vy + -2 * vy;
        if (m & 4) vz = vz + -2 * vz;
    }
}

__device__ __forceinline__ double dot3(double ax, double ay, double az, double bx,
                                       double by, double bz)
{
    // src/vec3_math.c:91-96: ((0 + a0 b0) + a1 b1) + a2 b2; 0 + v == v
    double o = ax * bx;
    o = o + ay * by;
    o = o + az * bz;
    return o;
}

// seg_pwrap, src/boundaries.c:54-83.  r2 moves by whole periods until it is
// within half a box of r1; the loops are capped (positions more than 64
// boxes away mean the state is already corrupt; the caller's finite-state
// check reports it).
__device__ __forceinline__ void pwrap(const TangleParams &P, double r1x, double r1y,
                                      double r1z, double &x, double &y, double &z)
{
    if (P.periodic[0]) {
        int g = 0;
        while ((x - r1x) > P.halfL[0] && g++ < 64) x -= P.L[0];
        while ((x - r1x) < -P.halfL[0] && g++ < 128) x += P.L[0];
    }
    if (P.periodic[1]) {
        int g = 0;
        while ((y - r1y) > P.halfL[1] && g++ < 64) y -= P.L[1];
        while ((y - r1y) < -P.halfL[1] && g++ < 128) y += P.L[1];
    }
    if (P.periodic[2]) {
        int g = 0;
        while ((z - r1z) > P.halfL[2] && g++ < 64) z -= P.L[2];
        while ((z - r1z) < -P.halfL[2] && g++ < 128) z += P.L[2];
    }
}

// The Biot-Savart field of a straight segment with end vectors R = r1 - r
// and Rp = r2 - r, src/tangle.c:540-559 (and the monopole of
// src/octree.c:327-342 with R = com - r, Rp = R + sum ds).  Reference
// guards: |R| or |Rp| < 1e-8, and the normalised dot (R.Rp/|R|)/|Rp| <
// 1e-8.  The two exact divisions of the second guard are only evaluated
// when a cheap bracket cannot decide it: q = dot / (|R| |Rp|) differs from
// the reference's (dot/|R|)/|Rp| by at most a few ulp, so a margin of
// 2^-40 relative decides every other case identically.
// Returns false when the pair contributes zero.
__device__ __forceinline__ bool seg_field64(double Rx, double Ry, double Rz, double Px,
                                            double Py, double Pz, double f_bs, double &vx,
                                            double &vy, double &vz)
{
    double lR = sqrt(dot3(Rx, Ry, Rz, Rx, Ry, Rz));
    double lP = sqrt(dot3(Px, Py, Pz, Px, Py, Pz));
    double dot = dot3(Rx, Ry, Rz, Px, Py, Pz);
    if (lR < 1e-8 || lP < 1e-8) return false;
    double ab = lR * lP;
    // bracket for ndot < 1e-8
    const double lo_t = 1e-8 * (1.0 - 0x1.0p-40), hi_t = 1e-8 * (1.0 + 0x1.0p-40);
    if (dot < lo_t * ab) return false;
    if (!(dot > hi_t * ab)) {
        double nd = dot / lR / lP;
        if (nd < 1e-8) return false;
    }
    double denom = ab * (ab + dot);
    double cx = Ry * Pz - Rz * Py;
    double cy = Rz * Px - Rx * Pz;
    double cz = Rx * Py - Ry * Px;
    double m = f_bs * (lR + lP) / denom;
    vx = cx * m;
    vy = cy * m;
    vz = cz * m;
    return true;
}

// 1/sqrt(x) and 1/x for normal x > 0 to double precision: the hardware
// estimates (about 2^-23 relative) refined by two Newton steps each
// (quadratic convergence: 2^-46, then below an ulp).
__device__ __forceinline__ double rsq64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = y * fma(-hx * y, y, 1.5);
    y = y * fma(-hx * y, y, 1.5);
    return y;
}

__device__ __forceinline__ double rcp64(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    y = y * fma(-x, y, 2.0);
    y = y * fma(-x, y, 2.0);
    return y;
}

// seg_field64 for the Barnes-Hut list traversal, where the pairs need the
// reference's guard decisions but not its rounding: 1/|R| and 1/|Rp| from
// rsq64 and the scalar from one rcp64 instead of two correctly rounded
// square roots and a division.  Every guard is decided as in seg_field64
// (|R|, |Rp| >= 1e-8 and (R.Rp/|R|)/|Rp| >= 1e-8): outside a 2^-40 band
// around each threshold the cheap values decide, inside it the exact
// expressions do.
__device__ __forceinline__ bool seg_field64_fast(double Rx, double Ry, double Rz, double Px,
                                                 double Py, double Pz, double f_bs,
                                                 double &vx, double &vy, double &vz)
{
    const double r2 = fma(Rz, Rz, fma(Ry, Ry, Rx * Rx));
    const double p2 = fma(Pz, Pz, fma(Py, Py, Px * Px));
    const double dot = fma(Rz, Pz, fma(Ry, Py, Rx * Px));
    const double t2 = 1e-16, lo = 1.0 - 0x1.0p-40, hi = 1.0 + 0x1.0p-40;
    if (r2 < t2 * lo || p2 < t2 * lo) return false;
    if (!(r2 > t2 * hi) && sqrt(dot3(Rx, Ry, Rz, Rx, Ry, Rz)) < 1e-8) return false;
    if (!(p2 > t2 * hi) && sqrt(dot3(Px, Py, Pz, Px, Py, Pz)) < 1e-8) return false;
    const double ia = rsq64(r2), ib = rsq64(p2);
    const double iab = ia * ib;
    const double nd = dot * iab;
    if (nd < 1e-8 * lo) return false;
    if (!(nd > 1e-8 * hi) &&
        dot3(Rx, Ry, Rz, Px, Py, Pz) / sqrt(dot3(Rx, Ry, Rz, Rx, Ry, Rz)) /
                sqrt(dot3(Px, Py, Pz, Px, Py, Pz)) <
            1e-8)
        return false;
    const double m = f_bs * (ia + ib) * iab * rcp64(1.0 + nd);
    vx = (Ry * Pz - Rz * Py) * m;
    vy = (Rz * Px - Rx * Pz) * m;
    vz = (Rx * Py - Ry * Px) * m;
    return true;
}

// A pair of the mixed-precision paths from FP64 R and Rp, in FP32 (native
// reciprocal square roots, one reciprocal, fused multiply-adds: this path
// makes no claim to the reference's rounding; the library is otherwise
// built with contraction off, and (|R|+|Rp|)/(|R||Rp|(|R||Rp| + R.Rp)) is
// evaluated as (ia + ib) ia ib / (1 + R.Rp ia ib) with ia = 1/|R|,
// ib = 1/|Rp|), except where FP32 cannot take the reference's guard
// decisions -- the normalised dot within 1e-5 of its threshold 1e-8 (FP32
// resolves it to about 1e-6), or |R|^2 or |Rp|^2 within a factor 2 of
// 1e-16 -- where the pair is evaluated by seg_field64, so every guard
// decides as in the reference (a flipped guard adds or drops the whole
// contribution).  Adds into (ax, ay, az); returns true when FP32 was used.
__device__ __forceinline__ bool seg_field_mixed(double Rx, double Ry, double Rz, double Px,
                                                double Py, double Pz, double f64, float f32,
                                                double &ax, double &ay, double &az)
{
    const float rx = (float)Rx, ry = (float)Ry, rz = (float)Rz;
    const float px = (float)Px, py = (float)Py, pz = (float)Pz;
    const float r2 = fmaf(rz, rz, fmaf(ry, ry, rx * rx));
    const float p2 = fmaf(pz, pz, fmaf(py, py, px * px));
    const float dot = fmaf(rz, pz, fmaf(ry, py, rx * px));
    const float ia = __builtin_amdgcn_rsqf(r2);
    const float ib = __builtin_amdgcn_rsqf(p2);
    const float iab = ia * ib;
    const float nd = dot * iab;
    if (fabsf(nd - 1e-8f) < 1e-5f || r2 < 2e-16f || p2 < 2e-16f) {
        double vx, vy, vz;
        if (seg_field64(Rx, Ry, Rz, Px, Py, Pz, f64, vx, vy, vz)) {
            ax += vx; ay += vy; az += vz;
        }
        return false;
    }
    if (nd < 1e-8f) return true;  // inside the Thales sphere: the reference adds 0
    const float m = f32 * (ia + ib) * iab * __builtin_amdgcn_rcpf(1.0f + nd);
    ax += (double)(fmaf(ry, pz, -(rz * py)) * m);
    ay += (double)(fmaf(rz, px, -(rx * pz)) * m);
    az += (double)(fmaf(rx, py, -(ry * px)) * m);
    return true;
}

// seg_field32 for far-field entries given R and the segment vector
// Q = Rp - R, without the guards, for pairs where none can trigger: the
// target lies outside the segment's Thales sphere (centre a + Q/2, radius
// |Q|/2) by a margin, so 1 + R.Rp/(|R||Rp|) stays away from 0 and |R|,
// |Rp| from 0 (the caller checks each entry against the whole target box
// and sends the others to the FP64 path, where the guards decide as in the
// reference).  Rp = R + Q is formed explicitly (a monopole's Q, the sum of
// its cell's segments, can reach past the target, where |Rp|^2 expanded as
// r2 + 2 R.Q + |Q|^2 would cancel).  R x Rp = R x Q, and with R = a - t (a
// the entry's and t the target's offset from a common centre)
// R x Q = a x Q - t x Q: the entry carries A = a x Q, the pair adds m A to
// sa and m Q to sq, and the caller forms sa - t x sq and applies f_bs once
// per batch of entries.  A zero entry placed far away (a = (1e15, 0, 0),
// Q = A = 0) adds exactly zero.
__device__ __forceinline__ void seg_field32_far_clear(float Rx, float Ry, float Rz, float Qx,
                                                      float Qy, float Qz, float Ax, float Ay,
                                                      float Az, float &sax, float &say,
                                                      float &saz, float &sqx, float &sqy,
                                                      float &sqz)
{
    const float Px = Rx + Qx, Py = Ry + Qy, Pz = Rz + Qz;
    const float r2 = fmaf(Rz, Rz, fmaf(Ry, Ry, Rx * Rx));
    const float p2 = fmaf(Pz, Pz, fmaf(Py, Py, Px * Px));
    const float dot = fmaf(Rz, Pz, fmaf(Ry, Py, Rx * Px));
    const float ia = __builtin_amdgcn_rsqf(r2);
    const float ib = __builtin_amdgcn_rsqf(p2);
    const float iab = ia * ib;
    const float m = (ia + ib) * iab * __builtin_amdgcn_rcpf(fmaf(dot, iab, 1.0f));
    sax = fmaf(Ax, m, sax);
    say = fmaf(Ay, m, say);
    saz = fmaf(Az, m, saz);
    sqx = fmaf(Qx, m, sqx);
    sqy = fmaf(Qy, m, sqy);
    sqz = fmaf(Qz, m, sqz);
}

}  // namespace ovg
