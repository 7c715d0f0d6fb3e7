"""The five benchmark workloads as seeded, synthetic ring tangles.

Each config is a periodic cube [0, D]^3 with the 26 periodic images of
src/boundary_images.c:45-72, kappa = 9.97e-4 cm^2/s and a0 = 1.3e-8 cm.
Reconnection distance is 0.8 * dl_min with dl_min = dxi/2 and
dl_max = 1.5 dxi, the reference defaults (src/initial_conditions.c:37-42)
scaled to the point spacing.  The counterflow v_ns is carried by the normal
fluid, v_n = v_ns x-hat with no imposed superfluid flow, and mutual friction
is on.
"""
import math

import openvort_amd as ov

KAPPA = 9.97e-4
A0 = 1.3e-8
T16 = (0.098, 0.016)   # 1.6 K: alpha, alpha'
T165 = (0.11, 0.014)   # 1.65 K
T19 = (0.206, 0.009)   # 1.9 K


def params(D, dxi, alpha_pair, v_ns, method=ov.METHOD_BH, precision=ov.PREC_FP64, theta=0.4,
           leaf_size=1, traverse=ov.TRAVERSE_POINT):
    p = ov.default_params(D)
    p.kappa = KAPPA
    p.vortex_width = A0
    p.alpha, p.alpha_p = alpha_pair
    p.use_mutual_friction = 1
    if v_ns:
        p.vn.type = ov.FLOW_SIMPLE
        p.vn.strength = v_ns
        p.vn.direction[0], p.vn.direction[1], p.vn.direction[2] = 1.0, 0.0, 0.0
    p.method = method
    p.precision = precision
    p.theta = theta
    p.leaf_size = leaf_size
    p.traverse = traverse
    p.rec_dist = 0.8 * (dxi / 2)
    p.rec_angle = 5 * math.pi / 180.0
    p.freq_to_cutoff = 4.0 / 5
    p.mixed_near = 2 * dxi
    return p


# name: (seed, rings, D, r_min, r_max, dxi, alpha pair, v_ns, dt)
CONFIGS = {
    "c1": (0xC0FFEE01, 16, 0.1, 0.01, 0.03, 1e-3, T16, 0.0, 1e-4),
    "c2": (0xC0FFEE02, 200, 0.1, 0.005, 0.03, 8e-4, T165, 1.0, 5e-5),
    "c3": (0xC0FFEE03, 4000, 0.2, 0.005, 0.05, 5e-4, T165, 1.5, 5e-5),
}


def make(name, rings=None, **kw):
    """(tangle, params, dt) of a named config; `rings` overrides the ring
    count (for reduced test sizes); keyword arguments go to params()."""
    seed, n_rings, D, rmin, rmax, dxi, ap, vns, dt = CONFIGS[name]
    if rings is not None:
        n_rings = rings
    t = ov.generate_rings(seed, n_rings, D, rmin, rmax, dxi)
    if name == "c1":
        kw.setdefault("method", ov.METHOD_DIRECT)
    p = params(D, dxi, ap, vns, **kw)
    return t, p, dt


def remesh_bounds(name):
    """(dl_min, dl_max) = (dxi/2, 1.5 dxi) of a config."""
    dxi = CONFIGS[name][5]
    return dxi / 2, 1.5 * dxi
