// ddgi_field.hpp — probe-field arithmetic of the irradiance probes: octahedral
// mapping, the probe rays' directions, atlas addressing and the lookup.  Host
// and device share it (the host uses it for nothing but checks of its own).
//
// Atlas layout: probe p owns the (side + 2)^2 tile at tile column p mod
// (count_x * count_y), tile row p div (count_x * count_y) (one tile row per
// z layer, as majercik2019_renderer.cpp:177-183 sizes the atlas).  Inside a
// tile, texels 1..side hold the octahedral map; the one-texel border holds
// copies of the texels across the map's folded edges, so bilinear filtering
// near an edge reads the right neighbours.
#pragma once

#include "rt_core.hpp"

namespace rt {

struct DdgiGrid {
    uint32_t count[3];
    float start[3], spacing[3];
    uint32_t side, tiles_x, atlas_w;  // tiles_x = count_x * count_y; atlas_w = tiles_x * (side + 2)
    uint32_t rays;                    // rays per probe
    float max_dist, sharpness, normal_bias, energy, hysteresis;
    uint32_t flags;                   // RT_DDGI_*
};

// Single-precision normalisation for the probe lookup, chosen for speed: a *
// (1 / sqrt(a . a)) with IEEE square root and division, the same on host and
// device.  GLSL defines normalize only as x / length(x), and drivers may use
// an approximate inverse square root, so whether this rounds as the
// reference's shaders do is not pinned.  The update's texel directions
// (oct_decode) keep the double normalize_ref: they are computed once per
// texel and update, off the lookup's hot path.
RT_HD f3 normalize_f(f3 a) {
    const float m = sqrtf(dot(a, a));
    if (m != 0.0f) {
        const float f = 1.0f / m;
        a = mk(a.x * f, a.y * f, a.z * f);
    }
    return a;
}

// octEncode (res/shader/helper.glsl:79-86)
RT_HD void oct_encode(f3 v, float& ox, float& oy) {
    const float l1 = fabsf(v.x) + fabsf(v.y) + fabsf(v.z);
    const float inv = 1.0f / l1;
    float rx = v.x * inv, ry = v.y * inv;
    if (v.z < 0.0f) {
        const float nx = (1.0f - fabsf(ry)) * (rx >= 0.0f ? 1.0f : -1.0f);
        const float ny = (1.0f - fabsf(rx)) * (ry >= 0.0f ? 1.0f : -1.0f);
        rx = nx;
        ry = ny;
    }
    ox = rx;
    oy = ry;
}

// octDecode (helper.glsl:88-94)
RT_HD f3 oct_decode(float ox, float oy) {
    f3 v = mk(ox, oy, 1.0f - fabsf(ox) - fabsf(oy));
    if (v.z < 0.0f) {
        const float nx = (1.0f - fabsf(oy)) * (ox >= 0.0f ? 1.0f : -1.0f);
        const float ny = (1.0f - fabsf(ox)) * (oy >= 0.0f ? 1.0f : -1.0f);
        v.x = nx;
        v.y = ny;
    }
    return normalize_ref(v);
}

// sphericalFibonacci(i, n) (helper.glsl:59-70): phi = 2 pi frac(i sqrt(5) / 2
// + 1/2 - 1), cos(theta) = 1 - (2 i + 1) / n.
RT_HD f3 fib_dir(uint32_t i, uint32_t n) {
    const float fi = (float)i;
    const float iphi = fi * 2.2360680f * 0.5f + 0.5f - 1.0f;
    float fr = iphi - floorf(iphi);
    if (!(fr < 1.0f)) fr = 0.0f;
    float s, c;
    sincos_2pi(fr, s, c);
    const float ct = 1.0f - (2.0f * fi + 1.0f) * (1.0f / (float)n);
    const float st = sqrtf(1.0f - ct * ct);
    return mk(c * st, s * st, ct);
}

// rot (row-major 3x3) * v, each row summed left to right
RT_HD f3 rot_apply(const float* r, f3 v) {
    return mk(r[0] * v.x + r[1] * v.y + r[2] * v.z, r[3] * v.x + r[4] * v.y + r[5] * v.z,
              r[6] * v.x + r[7] * v.y + r[8] * v.z);
}

// Probe position: x = p mod cx, y = (p div cx) mod cy, z = p div (cx cy)
// (majercik2019_probe_raygen.rgen:12-22), spacing * coord + start.
RT_HD f3 probe_pos(const DdgiGrid& g, uint32_t p) {
    const uint32_t x = p % g.count[0];
    const uint32_t y = (p / g.count[0]) % g.count[1];
    const uint32_t z = p / (g.count[0] * g.count[1]);
    return mk(g.spacing[0] * (float)x + g.start[0], g.spacing[1] * (float)y + g.start[1],
              g.spacing[2] * (float)z + g.start[2]);
}

// Interior texel (1..side, 1..side) whose value a tile texel (0..side+1)
// holds: itself inside, the mirrored texel across a folded edge on the
// border, the opposite corner at a corner.
RT_HD void tile_source(uint32_t side, uint32_t tx, uint32_t ty, uint32_t& ix, uint32_t& iy) {
    const uint32_t e = side + 1;
    const bool bx = tx == 0 || tx == e, by = ty == 0 || ty == e;
    if (bx && by) {
        ix = tx == 0 ? side : 1;
        iy = ty == 0 ? side : 1;
    } else if (bx) {
        ix = tx == 0 ? 1 : side;
        iy = e - ty;
    } else if (by) {
        ix = e - tx;
        iy = ty == 0 ? 1 : side;
    } else {
        ix = tx;
        iy = ty;
    }
}

// Direction of interior texel (ix, iy): the octahedral coordinate of its
// centre, (k + 0.5) * 2 / side - 1 (majercik2019_shading_update_comp.comp:15-21).
RT_HD f3 texel_dir(uint32_t side, uint32_t ix, uint32_t iy) {
    const float sc = 2.0f / (float)side;
    return oct_decode(((float)(ix - 1) + 0.5f) * sc - 1.0f, ((float)(iy - 1) + 0.5f) * sc - 1.0f);
}

RT_HD size_t tile_texel(const DdgiGrid& g, uint32_t p, uint32_t tx, uint32_t ty) {
    const uint32_t ts = g.side + 2;
    return (size_t)((p / g.tiles_x) * ts + ty) * g.atlas_w + (p % g.tiles_x) * ts + tx;
}

// Bilinear footprint of octahedral coordinate (ox, oy) in the tile at tile
// row trow, tile column tcol (probe (x, y, z) of the grid: z, x + y count_x):
// texel centres at k + 0.5, the map spanning tile coordinates [1, side + 1].
// Texel indices fit 32 bits (an atlas holds at most 2^28 texels).
struct Bilin {
    uint32_t t00, t10, t01, t11;
    float ax, ay;
};
RT_HD Bilin bilinear(const DdgiGrid& g, uint32_t trow, uint32_t tcol, float ox, float oy) {
    const float fs = (float)g.side;
    float fx = (ox + 1.0f) * 0.5f * fs + 0.5f;
    float fy = (oy + 1.0f) * 0.5f * fs + 0.5f;
    fx = fminf(fmaxf(fx, 0.0f), fs + 0.5f);  // [0.5, side + 0.5] for |o| <= 1; NaN maps to 0
    fy = fminf(fmaxf(fy, 0.0f), fs + 0.5f);
    const float x0 = floorf(fx), y0 = floorf(fy);
    Bilin b;
    b.ax = fx - x0;
    b.ay = fy - y0;
    const uint32_t ix = (uint32_t)x0, iy = (uint32_t)y0;
    const uint32_t ix1 = ix + 1 > g.side + 1 ? g.side + 1 : ix + 1;
    const uint32_t iy1 = iy + 1 > g.side + 1 ? g.side + 1 : iy + 1;
    const uint32_t ts = g.side + 2;
    const uint32_t r0 = (trow * ts + iy) * g.atlas_w + tcol * ts;
    const uint32_t r1 = (trow * ts + iy1) * g.atlas_w + tcol * ts;
    b.t00 = r0 + ix;
    b.t10 = r0 + ix1;
    b.t01 = r1 + ix;
    b.t11 = r1 + ix1;
    return b;
}
RT_HD float blend(const Bilin& b, float v00, float v10, float v01, float v11) {
    const float top = (1.0f - b.ax) * v00 + b.ax * v10;
    const float bot = (1.0f - b.ax) * v01 + b.ax * v11;
    return (1.0f - b.ay) * top + b.ay * bot;
}

// getIrradiance (majercik2019_final_raygen.rgen:43-91) at hit x with shading
// normal n, reached by a ray of direction vdir; energy_preservation applied.
// atl: two float4 per texel, (R, G, B, 0) then (mean, mean of squares, 0, 0),
// so the four texels of a bilinear footprint sit in two cache lines.
RT_HD f3 ddgi_lookup(const DdgiGrid& g, const float4* atl, f3 x, f3 n, f3 vdir) {
    const float xs[3] = {x.x, x.y, x.z};
    uint32_t bc[3];
    float alpha[3];
    for (int k = 0; k < 3; ++k) {
        // uvec3((X - start) / spacing) clamped to the grid (getBaseGridCoord)
        const float q = (xs[k] - g.start[k]) / g.spacing[k];
        const float top = (float)(g.count[k] - 1);
        bc[k] = q > 0.0f ? (uint32_t)fminf(q, top) : 0u;
        const float bp = g.spacing[k] * (float)bc[k] + g.start[k];
        alpha[k] = fminf(fmaxf(fabsf(xs[k] - bp) / g.spacing[k], 0.0f), 1.0f);
    }
    const f3 od = mk(-vdir.x, -vdir.y, -vdir.z);  // originDirection
    const f3 bias = scale(g.normal_bias, add(n, scale(3.0f, od)));
    const bool linear = (g.flags & RT_DDGI_LINEAR_BLENDING) != 0;
    const bool ref_vis = (g.flags & RT_DDGI_REFERENCE_VISIBILITY) != 0;
    float sr = 0.0f, sgr = 0.0f, sb = 0.0f, sw = 0.0f;
    for (uint32_t i = 0; i < 8; ++i) {
        const uint32_t off[3] = {i & 1u, (i >> 1) & 1u, (i >> 2) & 1u};
        uint32_t pc[3];
        for (int k = 0; k < 3; ++k) pc[k] = bc[k] + off[k] > g.count[k] - 1 ? g.count[k] - 1 : bc[k] + off[k];
        const f3 pp = mk(g.spacing[0] * (float)pc[0] + g.start[0], g.spacing[1] * (float)pc[1] + g.start[1],
                         g.spacing[2] * (float)pc[2] + g.start[2]);
        const f3 ptp = add(sub(x, pp), bias);  // probeToPoint
        const f3 dtp = normalize_f(mk(-ptp.x, -ptp.y, -ptp.z));
        float w = (dot(dtp, n) + 1.0f) * 0.5f;
        // lerp(1 - alpha, alpha, offset) per axis
        const float tx = off[0] ? alpha[0] : 1.0f - alpha[0];
        const float ty = off[1] ? alpha[1] : 1.0f - alpha[1];
        const float tz = off[2] ? alpha[2] : 1.0f - alpha[2];
        w = w * (tx * ty * tz);
        float ox, oy;
        oct_encode(normalize_f(mk(-dtp.x, -dtp.y, -dtp.z)), ox, oy);
        const Bilin b = bilinear(g, pc[2], pc[0] + pc[1] * g.count[0], ox, oy);
        const float dist = sqrtf(dot(ptp, ptp));
        const float4 i00 = atl[2 * b.t00], i10 = atl[2 * b.t10], i01 = atl[2 * b.t01], i11 = atl[2 * b.t11];
        const float4 d00 = atl[2 * b.t00 + 1], d10 = atl[2 * b.t10 + 1], d01 = atl[2 * b.t01 + 1],
                     d11 = atl[2 * b.t11 + 1];
        const float mean = blend(b, d00.x, d10.x, d01.x, d11.x);
        const float m2 = blend(b, d00.y, d10.y, d01.y, d11.y);
        const float var = fabsf(mean * mean - m2);
        const float dd = fmaxf(dist - mean, 0.0f);
        const float den = var + dd * dd;
        float cheb = den > 0.0f ? var / den : 1.0f;  // var = 0 at the mean: visible
        cheb = fmaxf(cheb * cheb * cheb, 0.0f);
        if (ref_vis ? dist <= mean : dist > mean) w = w * cheb;
        w = fmaxf(0.0001f, w);
        float ir = blend(b, i00.x, i10.x, i01.x, i11.x);
        float ig = blend(b, i00.y, i10.y, i01.y, i11.y);
        float ib = blend(b, i00.z, i10.z, i01.z, i11.z);
        if (!linear) {
            ir = sqrtf(fmaxf(ir, 0.0f));
            ig = sqrtf(fmaxf(ig, 0.0f));
            ib = sqrtf(fmaxf(ib, 0.0f));
        }
        sr = sr + w * ir;
        sgr = sgr + w * ig;
        sb = sb + w * ib;
        sw = sw + w;
    }
    f3 net = mk(sr / sw, sgr / sw, sb / sw);
    if (!linear) net = mulv(net, net);
    return scale(g.energy, net);
}

}  // namespace rt
