// Cheap path estimate of each ray and the surrogate features.
//
// Not a reference function: the cheap estimate whose residual against the
// full trace (tt_trace_rk4, the work of particle_tracker.py:312-380) the
// surrogate MLP learns.  The field is integrated along straight lines over
// z-blocks of k cells: block_kernel stores, per node column and block, the
// exact z-integral G of the trilinear field (the sum of the cells' mid-plane
// values) and its first moment M about the block centre, column by column.  Each ray then takes
// one drift-kick-drift step per block: drift half a block along its current
// direction, kick by G at that position, drift the second half.  Two
// first-moment terms make the straight-line integral follow the ray: the
// kick gets v . grad(M) (v = the ray's transverse drift in cells per cell),
// which is the integral of the field's linear variation along the tilted
// path, and the position gets -s M_x / W, because the kick is spread over the
// block instead of applied at its centre.  A ray entering inside the grid
// first takes single-cell steps from the node field up to the next block
// boundary.  For k = 1 the steps can read the node field directly (a cell's
// G and M follow from its two node planes): half the bytes per step of the
// block fields, and no block fields to build.  The full RK4 takes 8 field evaluations per z-cell (two steps of
// four stages); this takes one per k cells.
//
// Outputs per ray: base = [x, theta, y, phi, phase] (float64, the rf layout
// plus phase) and 48 float32 features, stored feature-major (feature q of ray
// r at feat[q * ldf + r], so every store of a wave is one contiguous run): 16 z-segment sums of the x kicks and of
// the y kicks (a step belongs to the segment of its first cell, 16 c / ncell),
// then U - U0, V - V0, W - W0, the exit displacement from the entry line (x,
// y, cells), the phase, U0, V0, W0 - 1, the entry-cell fractions in x and y,
// the numbers of column changes in x and y between steps, the summed
// second-order proxies q_x = (dGx/dx Gx + dGx/dy Gy) / W^2 and q_y = (dGy/dx
// Gx + dGy/dy Gy) / W^2 over the blocks, and the path length (cells).
#include "tt_internal.h"

namespace {

constexpr int kSeg = 16;

struct PathArgs {
    const float4 *field;  // node field [nz][ny][nx]
    const float4 *GM;     // per column and block: {G, M}, [ny][nx][nb][2] float4
    int k, nb;
    int nx, ny, nz;
    double x0, y0, z0, dx, dy, dz;
    float sx, sy;
    const double *s0;
    int64_t ld, n;
    const int32_t *order;
    float *feat;
    int64_t ldf;
    double *base;  // 5 x ldb
    int64_t ldb;
};

__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float t) {
    return make_float4(fmaf(t, b.x - a.x, a.x), fmaf(t, b.y - a.y, a.y), fmaf(t, b.z - a.z, a.z),
                       fmaf(t, b.w - a.w, a.w));
}
__device__ __forceinline__ float4 bilerp(const float4 *c, float tx, float ty) {
    return lerp4(lerp4(c[0], c[1], tx), lerp4(c[2], c[3], tx), ty);
}
__device__ __forceinline__ void load_corners(const float4 *p, int nx, float4 *c) {
    c[0] = p[0];
    c[1] = p[1];
    c[2] = p[nx];
    c[3] = p[nx + 1];
}

// atan for the small exit angles of a near-axial beam: below 1e-2 the
// series to x^7 is exact to double rounding (the next term, x^9 / 9, is
// below 1e-17 x); larger angles take the library atan.
__device__ __forceinline__ double small_atan(double x) {
    if (fabs(x) < 1e-2) {
        const double x2 = x * x;
        return x * fma(x2, fma(x2, fma(x2, -1.0 / 7.0, 1.0 / 5.0), -1.0 / 3.0), 1.0);
    }
    return atan(x);
}

struct PathState {
    float dpx, dpy, U, V, Wd;
    float iW;  // 1 / (1 + Wd), carried from the step that set Wd
    double ph;
    float segx, segy, crossx, crossy, qx, qy;
    int cix, ciy;
    int seg, seg_end;
};

// 1 / W by two Newton steps from a seed y near it
__device__ __forceinline__ float newton_rcp2(float W, float y) {
    float e = fmaf(-W, y, 1.0f);
    y = fmaf(y, e, y);
    e = fmaf(-W, y, 1.0f);
    return fmaf(y, e, y);
}

// Mid-step position of a step of width w from the current state (cells),
// its column (0 with inb = false outside the x/y range) and fractions.
struct Mid {
    float W, iW, dmx, dmy, tx, ty;
    int ix, iy;
    bool inb;
};

__device__ __forceinline__ Mid half_drift(const PathArgs &a, const PathState &S, float pxb, float pyb, float w) {
    Mid m;
    m.W = 1.0f + S.Wd;
    m.iW = S.iW;
    const float hm = 0.5f * w;
    m.dmx = fmaf(hm * a.sx * S.U, m.iW, S.dpx);
    m.dmy = fmaf(hm * a.sy * S.V, m.iW, S.dpy);
    const float mx = pxb + m.dmx, my = pyb + m.dmy;
    m.inb = mx >= 0.f && mx <= (float)(a.nx - 1) && my >= 0.f && my <= (float)(a.ny - 1);
    m.ix = m.inb ? min((int)mx, a.nx - 2) : 0;
    m.iy = m.inb ? min((int)my, a.ny - 2) : 0;
    m.tx = mx - (float)m.ix;
    m.ty = my - (float)m.iy;
    return m;
}

// Kick by g (the step's field integral), drift the second half with the
// position moment (mvx, mvy) and accumulate the features of a step that
// starts at cell c0.
__device__ __forceinline__ void kick_drift(const PathArgs &a, const Mid &m, float4 g, float w, float mvx, float mvy,
                                           int c0, int ncell, PathState &S, float *f) {
    const float hm = 0.5f * w;
    // W at the step's middle: g.z is the whole step's change of W (times W).
    // Its reciprocal, and the step end's, by two Newton steps each from the
    // nearby 1 / W (relative error e -> e^4: exact to fp32 rounding while W
    // changes by less than 1.5% over half a step) instead of v_rcp, which
    // costs several issue slots
    const float Wm = fmaf(0.5f * g.z, m.iW, m.W);
    float iWm = newton_rcp2(Wm, m.iW);
    const float sp = __builtin_amdgcn_sqrtf(fmaf(S.U, S.U, fmaf(S.V, S.V, m.W * m.W))) * m.iW;
    S.ph += (double)(sp * g.w);
    S.U = fmaf(g.x, iWm, S.U);
    S.V = fmaf(g.y, iWm, S.V);
    S.Wd = fmaf(g.z, iWm, S.Wd);
    const float iW1 = newton_rcp2(1.0f + S.Wd, iWm);
    S.iW = iW1;
    S.dpx = fmaf(fmaf(hm, S.U, -mvx) * a.sx, iW1, m.dmx);
    S.dpy = fmaf(fmaf(hm, S.V, -mvy) * a.sy, iW1, m.dmy);
    // z-segment sums: segment q holds cells [ceil(q ncell / 16), ceil((q + 1) ncell / 16))
    while (c0 >= S.seg_end) {
        f[S.seg * a.ldf] = S.segx;
        f[(kSeg + S.seg) * a.ldf] = S.segy;
        S.segx = S.segy = 0.f;
        ++S.seg;
        S.seg_end = ((S.seg + 1) * ncell + kSeg - 1) / kSeg;
    }
    S.segx += g.x;
    S.segy += g.y;
    if (S.cix >= 0) {
        S.crossx += (m.ix != S.cix) ? 1.f : 0.f;
        S.crossy += (m.iy != S.ciy) ? 1.f : 0.f;
    }
    S.cix = m.ix;
    S.ciy = m.iy;
}

constexpr float kTwelfth = 1.0f / 12.0f;

template <int J>
__device__ __forceinline__ float comp(const float4 &v) {
    return J == 0 ? v.x : J == 1 ? v.y : J == 2 ? v.z : v.w;
}

// Value and x / y slopes at (tx, ty) of the bilinear patch through corners
// c0 (0, 0), c1 (1, 0), c2 (0, 1), c3 (1, 1).
struct Patch {
    float v, dx, dy;
};
__device__ __forceinline__ Patch patch(float c0, float c1, float c2, float c3, float tx, float ty) {
    const float d01 = c1 - c0, d23 = c3 - c2;
    const float lo = fmaf(tx, d01, c0), hi = fmaf(tx, d23, c2);
    Patch p;
    p.dy = hi - lo;
    p.v = fmaf(ty, p.dy, lo);
    p.dx = fmaf(ty, d23 - d01, d01);
    return p;
}
template <int J>
__device__ __forceinline__ Patch patch4(const float4 (&c)[4], float tx, float ty) {
    return patch(comp<J>(c[0]), comp<J>(c[1]), comp<J>(c[2]), comp<J>(c[3]), tx, ty);
}

// One drift-kick-drift step over a block of w cells starting at cell b0,
// from the {G, M} values of the four columns around the mid-step position.
__device__ __forceinline__ void block_step(const PathArgs &a, const Mid &m, const float4 (&Gc)[4],
                                           const float4 (&Mc)[4], float w, int b0, int ncell, PathState &S, float *f) {
    const Patch g0 = patch4<0>(Gc, m.tx, m.ty), g1 = patch4<1>(Gc, m.tx, m.ty);
    const Patch g2 = patch4<2>(Gc, m.tx, m.ty), g3 = patch4<3>(Gc, m.tx, m.ty);
    const Patch m0 = patch4<0>(Mc, m.tx, m.ty), m1 = patch4<1>(Mc, m.tx, m.ty);
    const Patch m2 = patch4<2>(Mc, m.tx, m.ty), m3 = patch4<3>(Mc, m.tx, m.ty);
    // g = G + v . grad(M)
    const float vx = a.sx * S.U * m.iW, vy = a.sy * S.V * m.iW;
    float4 g = make_float4(fmaf(vy, m0.dy, fmaf(vx, m0.dx, g0.v)), fmaf(vy, m1.dy, fmaf(vx, m1.dx, g1.v)),
                           fmaf(vy, m2.dy, fmaf(vx, m2.dx, g2.v)), fmaf(vy, m3.dy, fmaf(vx, m3.dx, g3.v)));
    // second-order proxies (dGx/dx Gx + dGx/dy Gy) / W^2 and (dGy/dx Gx + dGy/dy Gy) / W^2
    const float iW2 = m.iW * m.iW;
    if (m.inb) {
        S.qx = fmaf(fmaf(g0.dx, g0.v, g0.dy * g1.v), iW2, S.qx);
        S.qy = fmaf(fmaf(g1.dx, g0.v, g1.dy * g1.v), iW2, S.qy);
    } else {
        g = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    kick_drift(a, m, g, w, m.inb ? m0.v : 0.f, m.inb ? m1.v : 0.f, b0, ncell, S, f);
}

// One node plane over a column square as its bilinear form
// f = A + B tx + C ty + E tx ty (corners c0 (0, 0), c1 (1, 0), c2 (0, 1),
// c3 (1, 1): B = c1 - c0, C = c2 - c0, E = c3 - c2 - B).
struct PlaneForm {
    float4 A, B, C, E;
};
__device__ __forceinline__ PlaneForm plane_form(const float4 *p, int nx) {
    const float4 c0 = p[0], c1 = p[1], c2 = p[nx], c3 = p[nx + 1];
    PlaneForm f;
    f.A = c0;
    f.B = make_float4(c1.x - c0.x, c1.y - c0.y, c1.z - c0.z, c1.w - c0.w);
    f.C = make_float4(c2.x - c0.x, c2.y - c0.y, c2.z - c0.z, c2.w - c0.w);
    f.E = make_float4(c3.x - c2.x - f.B.x, c3.y - c2.y - f.B.y, c3.z - c2.z - f.B.z, c3.w - c2.w - f.B.w);
    return f;
}
// value and x / y slopes of component J of a plane form at (tx, ty)
template <int J>
__device__ __forceinline__ Patch form_patch(const PlaneForm &f, float tx, float ty) {
    const float a = fmaf(tx, comp<J>(f.B), comp<J>(f.A));
    Patch p;
    p.dy = fmaf(tx, comp<J>(f.E), comp<J>(f.C));
    p.v = fmaf(ty, p.dy, a);
    p.dx = fmaf(ty, comp<J>(f.E), comp<J>(f.B));
    return p;
}

// One step over the single cell c from its node planes lo (c) and hi
// (c + 1) at the four columns around the mid-step position: the block step
// with G = (lo + hi) / 2 and M = (hi - lo) / 12 interpolated bilinearly
// (each plane interpolated once, the factors 1/2 and 1/12 applied to the
// interpolated sums and differences).
__device__ __forceinline__ void node_step(const PathArgs &a, const Mid &m, const PlaneForm &lo, const PlaneForm &hi,
                                          int c, int ncell, PathState &S, float *f) {
    const Patch l0 = form_patch<0>(lo, m.tx, m.ty), l1 = form_patch<1>(lo, m.tx, m.ty);
    const Patch l2 = form_patch<2>(lo, m.tx, m.ty), l3 = form_patch<3>(lo, m.tx, m.ty);
    const Patch h0 = form_patch<0>(hi, m.tx, m.ty), h1 = form_patch<1>(hi, m.tx, m.ty);
    const Patch h2 = form_patch<2>(hi, m.tx, m.ty), h3 = form_patch<3>(hi, m.tx, m.ty);
    // g = G + v . grad(M) = (s + (v / 6) . grad(d)) / 2, s = lo + hi, d = hi - lo
    const float vx6 = a.sx * S.U * m.iW * (1.0f / 6.0f), vy6 = a.sy * S.V * m.iW * (1.0f / 6.0f);
    float4 g = make_float4(0.5f * fmaf(vy6, h0.dy - l0.dy, fmaf(vx6, h0.dx - l0.dx, l0.v + h0.v)),
                           0.5f * fmaf(vy6, h1.dy - l1.dy, fmaf(vx6, h1.dx - l1.dx, l1.v + h1.v)),
                           0.5f * fmaf(vy6, h2.dy - l2.dy, fmaf(vx6, h2.dx - l2.dx, l2.v + h2.v)),
                           0.5f * fmaf(vy6, h3.dy - l3.dy, fmaf(vx6, h3.dx - l3.dx, l3.v + h3.v)));
    // second-order proxies (dGx/dx Gx + dGx/dy Gy) / W^2 with G = s / 2
    const float iW2q = 0.25f * m.iW * m.iW;
    const float s0v = l0.v + h0.v, s1v = l1.v + h1.v;
    if (m.inb) {
        S.qx = fmaf(fmaf(l0.dx + h0.dx, s0v, (l0.dy + h0.dy) * s1v), iW2q, S.qx);
        S.qy = fmaf(fmaf(l1.dx + h1.dx, s0v, (l1.dy + h1.dy) * s1v), iW2q, S.qy);
    } else {
        g = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    kick_drift(a, m, g, 1.0f, m.inb ? (h0.v - l0.v) * kTwelfth : 0.f, m.inb ? (h1.v - l1.v) * kTwelfth : 0.f, c,
               ncell, S, f);
}

// NODE: the k = 1 steps from the node field (no block fields); it runs at
// five waves per SIMD (a few registers spill outside the step loop), which
// hides more of each step's load latency; the block-field form keeps four.
template <bool NODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NODE ? 5 : 1))) void path_kernel(PathArgs a) {
    const int64_t t = (int64_t)tt::xcd_block((int)blockIdx.x, (int)gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const int64_t r = a.order ? (int64_t)a.order[t] : t;
    const double *s = a.s0 + r;
    const double x = s[0], y = s[a.ld], z = s[2 * a.ld];
    const double U0 = s[3 * a.ld] / tt::kC, V0 = s[4 * a.ld] / tt::kC, W0 = s[5 * a.ld] / tt::kC;
    float *f = a.feat + r;  // feature-major: feature q of this ray at f[q * ldf]
    double *b = a.base + r;
    const double nanv = __longlong_as_double(0x7ff8000000000000ll);
    // a ray the full trace hands over at launch (W0 <= 1 + kWdMin, which
    // includes W0 <= 0) is out of the estimate's reach: NaN, so the surrogate
    // flags it for the fallback
    if (!(W0 - 1.0 > (double)tt::kWdMin) || isnan(x) || isnan(y) || isnan(z) || isnan(U0) || isnan(V0)) {
        for (int c = 0; c < 5; ++c) b[c * a.ldb] = nanv;
        for (int c = 0; c < 48; ++c) f[c * a.ldf] = 0.f;
        return;
    }
    double px0 = (x - a.x0) / a.dx, py0 = (y - a.y0) / a.dy, Zs = (z - a.z0) / a.dz;
    if (Zs < 0.0) {  // zero field before the grid: straight to the first plane
        px0 += (double)a.sx * U0 / W0 * -Zs;
        py0 += (double)a.sy * V0 / W0 * -Zs;
        Zs = 0.0;
    }
    const int ncell = a.nz - 1;
    const float pxb = (float)px0, pyb = (float)py0;
    PathState S;
    S.dpx = S.dpy = 0.f;
    S.U = (float)U0;
    S.V = (float)V0;
    S.Wd = (float)(W0 - 1.0);
    S.iW = __builtin_amdgcn_rcpf(1.0f + S.Wd);
    S.ph = 0.0;
    S.segx = S.segy = S.crossx = S.crossy = S.qx = S.qy = 0.f;
    S.cix = S.ciy = -1;
    for (int q = 0; q < 2 * kSeg; ++q) f[q * a.ldf] = 0.f;
    const int iz0 = Zs >= (double)ncell ? ncell : (int)Zs;
    S.seg = (iz0 * kSeg) / (ncell > 0 ? ncell : 1);
    S.seg_end = ((S.seg + 1) * ncell + kSeg - 1) / kSeg;
    const bool on_boundary = Zs == (double)iz0 && iz0 % a.k == 0;
    const int bstart = on_boundary ? iz0 : min(ncell, (iz0 / a.k + 1) * a.k);
    // single-cell steps from the node field: the first one may be partial,
    // sampled at its mid-step plane 1 - w/2 between planes c and c + 1
    for (int c = iz0; c < bstart; ++c) {
        const float w = c == iz0 ? (float)((double)(c + 1) - Zs) : 1.0f;
        const Mid m = half_drift(a, S, pxb, pyb, w);
        const float4 *p = a.field + ((int64_t)c * a.ny + m.iy) * a.nx + m.ix;
        float4 lo[4], hi[4];
        load_corners(p, a.nx, lo);
        load_corners(p + (int64_t)a.nx * a.ny, a.nx, hi);
        const float tz = 1.0f - 0.5f * w;
        for (int q = 0; q < 4; ++q) lo[q] = lerp4(lo[q], hi[q], tz);
        float4 g = bilerp(lo, m.tx, m.ty);
        g = m.inb ? make_float4(w * g.x, w * g.y, w * g.z, w * g.w) : make_float4(0.f, 0.f, 0.f, 0.f);
        kick_drift(a, m, g, w, 0.f, 0.f, c, ncell, S, f);
    }
    if (NODE) {
        // k = 1 without block fields: a cell's G and M follow from its two
        // node planes, G = (lo + hi) / 2 and M = (hi - lo) / 12 (what
        // block_kernel stores for k = 1), so each step reads the node field:
        // the upper plane's four corners, and the lower plane's only when the
        // ray has changed column (else it is the last step's upper plane, kept
        // as its bilinear form).  Half the bytes per step of the block fields,
        // no block fields to build; loading the next upper plane one step
        // ahead was slower (more registers, fewer waves).
        const int64_t plane = (int64_t)a.nx * a.ny;
        int qx = -1, qy = -1;
        PlaneForm lo;
        for (int c = bstart; c < ncell; ++c) {
            const Mid m = half_drift(a, S, pxb, pyb, 1.0f);
            const float4 *p = a.field + ((int64_t)(c + 1) * a.ny + m.iy) * a.nx + m.ix;
            const PlaneForm hi = plane_form(p, a.nx);
            if (m.ix != qx || m.iy != qy) {
                lo = plane_form(p - plane, a.nx);
                qx = m.ix;
                qy = m.iy;
            }
            node_step(a, m, lo, hi, c, ncell, S, f);
            lo = hi;
        }
    } else {
        // z-blocks: a column's blocks are consecutive in memory, so a wave's
        // successive steps read neighbouring bytes of the same few columns
        const int64_t cstride = 2 * (int64_t)a.nb;  // float4 per column
        for (int bb = bstart / a.k; bb < a.nb && bstart < ncell; ++bb) {
            const int b0 = bb * a.k;
            const float w = (float)(min(ncell, b0 + a.k) - b0);
            const Mid m = half_drift(a, S, pxb, pyb, w);
            const float4 *q = a.GM + ((int64_t)m.iy * a.nx + m.ix) * cstride + 2 * bb;
            float4 Gc[4], Mc[4];
            Gc[0] = q[0];
            Mc[0] = q[1];
            Gc[1] = q[cstride];
            Mc[1] = q[cstride + 1];
            Gc[2] = q[a.nx * cstride];
            Mc[2] = q[a.nx * cstride + 1];
            Gc[3] = q[(a.nx + 1) * cstride];
            Mc[3] = q[(a.nx + 1) * cstride + 1];
            block_step(a, m, Gc, Mc, w, b0, ncell, S, f);
        }
    }
    if (S.seg < kSeg) {
        f[S.seg * a.ldf] = S.segx;
        f[(kSeg + S.seg) * a.ldf] = S.segy;
    }
    const float U = S.U, V = S.V, Wd = S.Wd, dpx = S.dpx, dpy = S.dpy;
    const double ph = S.ph;
    const double Uf = U, Vf = V, Wf = 1.0 + (double)Wd;
    const double pxe = px0 + (double)dpx, pye = py0 + (double)dpy;
    // the full trace's other hand-overs, judged on the estimate's own path:
    // W at or below 1 + kWdMin at the exit, or a side face of the grid crossed
    // (x / y inside the grid at the exit but not at the entry, or the reverse)
    const double fx1 = (double)(a.nx - 1), fy1 = (double)(a.ny - 1);
    const bool in0 = px0 >= 0.0 && px0 <= fx1 && py0 >= 0.0 && py0 <= fy1;
    const bool in1 = pxe >= 0.0 && pxe <= fx1 && pye >= 0.0 && pye <= fy1;
    const bool handover = !(Wd > tt::kWdMin) || in0 != in1;
    b[0] = handover ? nanv : a.x0 + pxe * a.dx;
    b[a.ldb] = handover ? nanv : small_atan(Uf / Wf);
    b[2 * a.ldb] = handover ? nanv : a.y0 + pye * a.dy;
    b[3 * a.ldb] = handover ? nanv : small_atan(Vf / Wf);
    b[4 * a.ldb] = handover ? nanv : ph;
    const double L = (double)ncell - Zs;
    float *gf = f + 2 * kSeg * a.ldf;
    const int64_t L1 = a.ldf;
    gf[0 * L1] = (float)(Uf - U0);
    gf[1 * L1] = (float)(Vf - V0);
    gf[2 * L1] = Wd - (float)(W0 - 1.0);
    gf[3 * L1] = (float)(pxe - (px0 + (double)a.sx * U0 / W0 * L));
    gf[4 * L1] = (float)(pye - (py0 + (double)a.sy * V0 / W0 * L));
    gf[5 * L1] = (float)ph;
    gf[6 * L1] = (float)U0;
    gf[7 * L1] = (float)V0;
    gf[8 * L1] = (float)(W0 - 1.0);
    gf[9 * L1] = (float)(px0 - floor(px0));
    gf[10 * L1] = (float)(py0 - floor(py0));
    gf[11 * L1] = S.crossx;
    gf[12 * L1] = S.crossy;
    gf[13 * L1] = S.qx;
    gf[14 * L1] = S.qy;
    gf[15 * L1] = (float)L;
}

// Block integrals: for each block of cells [b0, b1) of a node column,
// G = sum_c (f[c] + f[c+1]) / 2 and
// M = sum_c (c + 1/2 - zmid) (f[c] + f[c+1]) / 2 + (f[b1] - f[b0]) / 12,
// stored as GM[col][b] = {G, M}.  Thread per (node column, block): blockIdx.y
// is the block, so the grid has plane x nb threads and every thread streams
// only its block's k + 1 planes.
__global__ __launch_bounds__(256) void block_kernel(const float4 *__restrict__ f, float4 *__restrict__ GM,
                                                    int64_t plane, int ncell, int k, int nb) {
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= plane) return;
    const int bb = blockIdx.y;
    const int b0 = bb * k, b1 = min(ncell, b0 + k);
    const float zm = 0.5f * (float)(b0 + b1);
    float4 lo = f[(int64_t)b0 * plane + col];
    const float4 first = lo;
    float4 g = make_float4(0.f, 0.f, 0.f, 0.f), m = g;
    for (int c = b0; c < b1; ++c) {
        const float4 hi = f[(int64_t)(c + 1) * plane + col];
        const float4 mid = make_float4(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z),
                                       0.5f * (lo.w + hi.w));
        const float zc = (float)c + 0.5f - zm;
        g.x += mid.x;
        g.y += mid.y;
        g.z += mid.z;
        g.w += mid.w;
        m.x = fmaf(zc, mid.x, m.x);
        m.y = fmaf(zc, mid.y, m.y);
        m.z = fmaf(zc, mid.z, m.z);
        m.w = fmaf(zc, mid.w, m.w);
        lo = hi;
    }
    m.x = fmaf(lo.x - first.x, kTwelfth, m.x);
    m.y = fmaf(lo.y - first.y, kTwelfth, m.y);
    m.z = fmaf(lo.z - first.z, kTwelfth, m.z);
    m.w = fmaf(lo.w - first.w, kTwelfth, m.w);
    float4 *out = GM + (col * nb + bb) * 2;
    out[0] = g;
    out[1] = m;
}

// corrected = base + Y[:, :5] * scale; flag when a predicted error bound
// (|Y[:, 5]| * e_pos for position, |Y[:, 6]| * e_ang for angle) exceeds its
// threshold, or when the estimate is not finite.  Flagged ray indices, plus
// `first` (the index of this batch's first ray in a larger ray set), are
// appended to list (count in *nflag).
__global__ void apply_kernel(const double *base, int64_t ldb, const float *Y, int64_t ldy, int64_t n, float s_pos,
                             float s_ang, float s_ph, float e_pos, float e_ang, double t_pos, double t_ang,
                             double *rf, int64_t ldr, double *phase, int32_t *flag, int32_t *list, int32_t *nflag,
                             int64_t first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;  // the rest of the wave stays converged below
    const float *y = Y + i * ldy;
    const double x = base[i] + (double)(y[0] * s_pos);
    const double th = base[ldb + i] + (double)(y[2] * s_ang);
    const double yy = base[2 * ldb + i] + (double)(y[1] * s_pos);
    const double ph = base[3 * ldb + i] + (double)(y[3] * s_ang);
    const double pz = base[4 * ldb + i] + (double)(y[4] * s_ph);
    rf[i] = x;
    rf[ldr + i] = th;
    rf[2 * ldr + i] = yy;
    rf[3 * ldr + i] = ph;
    if (phase) phase[i] = pz;
    const double ep = fabs((double)(y[5] * e_pos)), ea = fabs((double)(y[6] * e_ang));
    const bool bad = !(ep <= t_pos && ea <= t_ang) || isnan(x) || isnan(th) || isnan(yy) || isnan(ph);
    if (flag) flag[i] = bad ? 1 : 0;
    if (list) {
        // one atomic per wave: the wave's flagged rays take consecutive
        // slots in lane order, so the list keeps the rays' entry-column
        // order within each wave (the fallback trace stays coherent)
        const uint64_t m = __ballot(bad);
        if (m) {
            const int lane = (int)__lane_id();
            const int leader = __ffsll((unsigned long long)m) - 1;
            int slot = 0;
            if (lane == leader) slot = atomicAdd(nflag, (int)__popcll(m));
            slot = __shfl(slot, leader, 64);
            if (bad) list[slot + (int)__popcll(m & ((1ull << lane) - 1ull))] = (int32_t)(first + i);
        }
    }
}

}  // namespace

extern "C" int tt_surrogate_apply(const double *base, int64_t ldb, const float *Y, int64_t ldy, int64_t n,
                                  const float *scales, const double *thresholds, double *rf, int64_t ldr,
                                  double *phase, int32_t *flag, int32_t *list, int32_t *nflag, int64_t first,
                                  void *stream) {
    TT_REQUIRE(n >= 0 && ldb >= n && ldr >= n && ldy >= 7, "tt_surrogate_apply: bad sizes");
    TT_REQUIRE(first >= 0 && first + n <= 0x7fffffff, "tt_surrogate_apply: ray indices must fit int32");
    if (n == 0) return TT_OK;  // an empty ray set: nothing is read or written
    TT_REQUIRE(base && Y && scales && thresholds && rf, "tt_surrogate_apply: null pointer");
    TT_REQUIRE(!list || nflag, "tt_surrogate_apply: list needs nflag");
    hipLaunchKernelGGL(apply_kernel, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0, tt::as_stream(stream), base,
                       ldb, Y, ldy, n, scales[0], scales[1], scales[2], scales[3], scales[4], thresholds[0],
                       thresholds[1], rf, ldr, phase, flag, list, nflag, first);
    TT_LAUNCH_CHECK("tt_surrogate_apply launch");
    return TT_OK;
}

extern "C" int tt_field_blocks(const float *field4, const tt_grid *grid, int32_t k, float *GM4, void *stream) {
    TT_REQUIRE(field4 && grid && GM4, "tt_field_blocks: null pointer");
    TT_REQUIRE(grid->nx >= 2 && grid->ny >= 2 && grid->nz >= 2, "tt_field_blocks: grid needs >= 2 nodes per axis");
    TT_REQUIRE(k >= 1 && k <= 1024, "tt_field_blocks: block size k must be in [1, 1024]");
    const int ncell = grid->nz - 1, nb = (ncell + k - 1) / k;
    TT_REQUIRE(nb <= 65535, "tt_field_blocks: too many blocks");
    const int64_t plane = (int64_t)grid->nx * grid->ny;
    hipLaunchKernelGGL(block_kernel, dim3((unsigned)tt::ceil_div(plane, 256), (unsigned)nb), dim3(256), 0,
                       tt::as_stream(stream),
                       reinterpret_cast<const float4 *>(field4), reinterpret_cast<float4 *>(GM4), plane, ncell, (int)k,
                       nb);
    TT_LAUNCH_CHECK("tt_field_blocks launch");
    return TT_OK;
}

extern "C" int tt_path_features(const float *field4, const float *GM4, int32_t k, const tt_grid *grid, const double *s0,
                                int64_t ld, int64_t n, const int32_t *order, float *feat, int64_t ldf, double *base,
                                int64_t ldb, void *stream) {
    TT_REQUIRE(grid, "tt_path_features: null grid");
    TT_REQUIRE(n >= 0 && ld >= n && ldb >= n && ldf >= n, "tt_path_features: bad sizes");
    TT_REQUIRE(grid->nx >= 2 && grid->ny >= 2 && grid->nz >= 2, "tt_path_features: grid needs >= 2 nodes per axis");
    TT_REQUIRE(grid->dx > 0 && grid->dy > 0 && grid->dz > 0, "tt_path_features: spacing must be > 0");
    TT_REQUIRE(k >= 1 && k <= 1024, "tt_path_features: block size k must be in [1, 1024]");
    if (n == 0) return TT_OK;  // an empty ray set: nothing is read or written
    TT_REQUIRE(field4 && s0 && feat && base, "tt_path_features: null pointer");
    TT_REQUIRE(GM4 || k == 1, "tt_path_features: block fields GM4 are required for k > 1");
    PathArgs a;
    a.field = reinterpret_cast<const float4 *>(field4);
    a.GM = reinterpret_cast<const float4 *>(GM4);
    a.k = k;
    a.nb = (grid->nz - 1 + k - 1) / k;
    a.nx = grid->nx;
    a.ny = grid->ny;
    a.nz = grid->nz;
    a.x0 = grid->x0;
    a.y0 = grid->y0;
    a.z0 = grid->z0;
    a.dx = grid->dx;
    a.dy = grid->dy;
    a.dz = grid->dz;
    a.sx = (float)(grid->dz / grid->dx);
    a.sy = (float)(grid->dz / grid->dy);
    a.s0 = s0;
    a.ld = ld;
    a.n = n;
    a.order = order;
    a.feat = feat;
    a.ldf = ldf;
    a.base = base;
    a.ldb = ldb;
    if (a.GM)
        hipLaunchKernelGGL(path_kernel<false>, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0,
                           tt::as_stream(stream), a);
    else
        hipLaunchKernelGGL(path_kernel<true>, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0,
                           tt::as_stream(stream), a);
    TT_LAUNCH_CHECK("tt_path_features launch");
    return TT_OK;
}
