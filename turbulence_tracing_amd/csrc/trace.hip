// Ray integrator: the work of ElectronCube.solve + ray_at_exit and of every
// dsdt/dndr evaluation inside solve_ivp
// (particle_tracking/particle_tracker.py:243-256, 312-380, 398-419).
//
// The reference integrates d x/dt = v, d v/dt = -(c^2/2) grad(ne/nc) in
// time with an adaptive RK45 shared by all rays.  Here each ray is one lane
// and the same ODE is integrated with z (in cells) as the independent
// variable by classic RK4 with a fixed step of h = 1/m cells:
//   d px/dZ = (dz/dx) U/W   d py/dZ = (dz/dy) V/W
//   d U/dZ  = Gx / W        d V/dZ  = Gy / W        d W/dZ = Gz / W
//   d tau/dZ = dz / W       d phase/dZ = Phi |v| / W
// (U, V, W) = v/c and (Gx, Gy, Gz, Phi) is the trilinear interpolant of the
// float4 node field written by tt_grad_stencil (zero strictly outside the
// grid, like RegularGridInterpolator with fill_value=0).  Steps land on the
// z-planes, so each step stays inside one z-cell and the last lands exactly
// on the exit plane z = z.max() that ray_at_exit back-projects to.
//
// Precision: positions are carried as an fp64 base plus an fp32 displacement
// from the entry point, W as W - 1, and the phase is summed per step into an
// fp64 accumulator.
//
// Algorithmic traffic (roofline accounting): 4 stages x 8 corners x 16 B =
// 512 B per ray-step.
#include <algorithm>
#include <vector>

#include "tt_internal.h"

namespace {

struct TraceArgs {
    const float4 *field;
    const float4 *forms;  // optional plane forms (tt_plane_forms): 4 float4 per node
    int nx, ny, nz;
    uint32_t plane;       // nx ny (< 2^31)
    float fx1, fy1, fz1;  // n - 1 as float (upper bounds, inclusive)
    double x0, y0, z0, dx, dy, dz;
    float sx, sy;         // dz/dx, dz/dy
    int m;                // steps per cell
    float h;              // 1/m
    double cT;
    const double *s0;
    int64_t ld_s0, n;
    const int32_t *order;
    double *rf;
    int64_t ld_rf;
    double *phase;
    double *sf;
    int64_t ld_sf;
    int32_t *status;
    const int32_t *n_dev;  // optional (device): trace only launch indices below *n_dev
};

// qx, qy: (x, y) displacement from the entry point in units of dz (the
// displacement in cells is (dz/dx) qx, (dz/dy) qy); Wd = W - 1;
// tr = int (1/W - 1) dZ
struct State {
    float qx, qy, U, V, Wd, tr;
};

__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float t) {
    return make_float4(fmaf(t, b.x - a.x, a.x), fmaf(t, b.y - a.y, a.y), fmaf(t, b.z - a.z, a.z),
                       fmaf(t, b.w - a.w, a.w));
}

// Index of node (ix, iy, iz), all >= 0: iz (nx ny) + (iy nx + ix) as one
// 32 x 32 -> 64-bit multiply-add and one 24-bit multiply-add (nx, ny < 2^24,
// nx ny < 2^31), instead of a 64-bit multiply by nx.
__device__ __forceinline__ int64_t node_index(const TraceArgs &a, int ix, int iy, int iz) {
    const uint32_t row = __umul24((uint32_t)iy, (uint32_t)a.nx) + (uint32_t)ix;
    return (int64_t)((uint64_t)(uint32_t)iz * a.plane + row);
}

// Trilinear sample in cell coordinates; zero strictly outside [0, n-1]^3.
__device__ __forceinline__ float4 sample(const TraceArgs &a, float px, float py, float pz) {
    if (!(px >= 0.f && px <= a.fx1 && py >= 0.f && py <= a.fy1 && pz >= 0.f && pz <= a.fz1))
        return make_float4(0.f, 0.f, 0.f, 0.f);
    const int ix = min((int)px, a.nx - 2);
    const int iy = min((int)py, a.ny - 2);
    const int iz = min((int)pz, a.nz - 2);
    const float tx = px - (float)ix, ty = py - (float)iy, tz = pz - (float)iz;
    const int64_t plane = (int64_t)a.nx * a.ny;
    const float4 *p = a.field + node_index(a, ix, iy, iz);
    const float4 c000 = p[0], c100 = p[1], c010 = p[a.nx], c110 = p[a.nx + 1];
    const float4 c001 = p[plane], c101 = p[plane + 1], c011 = p[plane + a.nx], c111 = p[plane + a.nx + 1];
    const float4 c00 = lerp4(c000, c100, tx), c10 = lerp4(c010, c110, tx);
    const float4 c01 = lerp4(c001, c101, tx), c11 = lerp4(c011, c111, tx);
    return lerp4(lerp4(c00, c10, ty), lerp4(c01, c11, ty), tz);
}

// Path-length factor ds/(dz dZ) = |v| / W of a state.  The phase slope of a
// stage is Phi |v| / W; the RK4 steps take the field part Phi at the four
// stage points and |v| / W once, at the third stage's state (the step's
// midpoint): |v| / W = sqrt(1 + tan^2) changes by O(angle x deflection) per
// cell, so the midpoint value differs from per-stage values by far less
// than the fp32 rounding of the increment.
__device__ __forceinline__ float path_factor(const State &s, float W, float iW) {
    return __builtin_amdgcn_sqrtf(fmaf(s.U, s.U, fmaf(s.V, s.V, W * W))) * iW;
}

template <bool VF>
__device__ __forceinline__ void rhs(const TraceArgs &a, float pxb, float pyb, float Z, const State &s, State &k,
                                    float &gw, float &vf) {
    const float4 g = sample(a, fmaf(a.sx, s.qx, pxb), fmaf(a.sy, s.qy, pyb), Z);
    const float W = 1.0f + s.Wd;
    const float iW = __builtin_amdgcn_rcpf(W);
    k.qx = s.U * iW;
    k.qy = s.V * iW;
    k.U = g.x * iW;
    k.V = g.y * iW;
    k.Wd = g.z * iW;
    k.tr = -s.Wd * iW;
    gw = g.w;
    if (VF) vf = path_factor(s, W, iW);
}

__device__ __forceinline__ State axpy(const State &y, float c, const State &k) {
    return State{fmaf(c, k.qx, y.qx), fmaf(c, k.qy, y.qy), fmaf(c, k.U, y.U),
                 fmaf(c, k.V, y.V),   fmaf(c, k.Wd, y.Wd), fmaf(c, k.tr, y.tr)};
}

// one classic RK4 step of size hh from Z; returns the phase increment
__device__ __forceinline__ float rk4_step(const TraceArgs &a, float pxb, float pyb, float Z, float hh, State &y) {
    State k1, k2, k3, k4;
    float p1, p2, p3, p4, vf = 0.f;
    const float h2 = 0.5f * hh;
    rhs<false>(a, pxb, pyb, Z, y, k1, p1, vf);
    rhs<false>(a, pxb, pyb, Z + h2, axpy(y, h2, k1), k2, p2, vf);
    rhs<true>(a, pxb, pyb, Z + h2, axpy(y, h2, k2), k3, p3, vf);
    rhs<false>(a, pxb, pyb, Z + hh, axpy(y, hh, k3), k4, p4, vf);
    const float h6 = hh * (1.0f / 6.0f);
    y.qx += h6 * (k1.qx + 2.f * (k2.qx + k3.qx) + k4.qx);
    y.qy += h6 * (k1.qy + 2.f * (k2.qy + k3.qy) + k4.qy);
    y.U += h6 * (k1.U + 2.f * (k2.U + k3.U) + k4.U);
    y.V += h6 * (k1.V + 2.f * (k2.V + k3.V) + k4.V);
    y.Wd += h6 * (k1.Wd + 2.f * (k2.Wd + k3.Wd) + k4.Wd);
    y.tr += h6 * (k1.tr + 2.f * (k2.tr + k3.tr) + k4.tr);
    return h6 * (p1 + 2.f * (p2 + p3) + p4) * vf;
}

// One classic RK4 step as rk4_step, returning the phase integrand
// Phi |v| / W at the step's start point and, in dph, the step's phase by the
// RK4 weights (the h = 1/2 march uses one or the other).
__device__ __forceinline__ float rk4_step_phi(const TraceArgs &a, float pxb, float pyb, float Z, float hh, State &y,
                                              float &dph) {
    State k1, k2, k3, k4;
    float p1, p2, p3, p4, vf1 = 0.f, vf = 0.f, unused;
    const float h2 = 0.5f * hh;
    rhs<true>(a, pxb, pyb, Z, y, k1, p1, vf1);
    rhs<false>(a, pxb, pyb, Z + h2, axpy(y, h2, k1), k2, p2, unused);
    rhs<true>(a, pxb, pyb, Z + h2, axpy(y, h2, k2), k3, p3, vf);
    rhs<false>(a, pxb, pyb, Z + hh, axpy(y, hh, k3), k4, p4, unused);
    const float h6 = hh * (1.0f / 6.0f);
    y.qx += h6 * (k1.qx + 2.f * (k2.qx + k3.qx) + k4.qx);
    y.qy += h6 * (k1.qy + 2.f * (k2.qy + k3.qy) + k4.qy);
    y.U += h6 * (k1.U + 2.f * (k2.U + k3.U) + k4.U);
    y.V += h6 * (k1.V + 2.f * (k2.V + k3.V) + k4.V);
    y.Wd += h6 * (k1.Wd + 2.f * (k2.Wd + k3.Wd) + k4.Wd);
    y.tr += h6 * (k1.tr + 2.f * (k2.tr + k3.tr) + k4.tr);
    dph = h6 * (p1 + 2.f * (p2 + p3) + p4) * vf;
    return p1 * vf1;
}

// The phase integrand Phi |v| / W by trilinear gathers at plane Z (cells),
// at the position of state s moved back by `back` cells of z along its
// direction (used at the ends of the march and, in the rare cells where one
// step crossed a face, at the other step's midpoint).
__device__ __forceinline__ float phi_at(const TraceArgs &a, float pxb, float pyb, float Z, const State &s,
                                        float back = 0.f) {
    const float W = 1.0f + s.Wd;
    const float iW = __builtin_amdgcn_rcpf(W);
    const float qx = fmaf(-back * s.U, iW, s.qx), qy = fmaf(-back * s.V, iW, s.qy);
    const float4 g = sample(a, fmaf(a.sx, qx, pxb), fmaf(a.sy, qy, pyb), Z);
    return g.w * path_factor(s, W, iW);
}

// Hand-over rule of the z-march.  A ray leaves the march (status 1, traced
// in time by tt_trace_time instead) when, at launch or at the end of any
// step, W = vz / c <= 0.97 (a slope sqrt(U^2 + V^2) / W above ~0.25, or a ray
// slowing in z towards a turning point), or when it crosses a side face of
// the grid (x or y), where the field drops to zero and a fixed step across
// the jump loses its order.  Below that W the fixed half-cell z-step is no
// longer accurate to 1e-3 cells, and near W = 0 it can step over a turning
// point.  With ne/nc clipped at 1 the field changes W by at most ~0.3 per
// step, so a ray cannot pass from W > 0.97 to a turning point within one
// step unnoticed.  The test costs nothing beyond the W > 0 test it replaces.
// The path estimate (path.hip) flags the same rays for the surrogate's
// fallback.
using tt::kWdMin;  // W - 1 at the hand-over

__device__ __forceinline__ bool lateral_inside(const TraceArgs &a, float pxb, float pyb, const State &s) {
    const float px = fmaf(a.sx, s.qx, pxb), py = fmaf(a.sy, s.qy, pyb);
    return px >= 0.f && px <= a.fx1 && py >= 0.f && py <= a.fy1;
}

// ---------------------------------------------------------------------------
// Cell-cached evaluation.  A lane loads the cell its ray is in once, as the
// trilinear polynomial f = a + b tx + c ty + d tz + e tx ty + f tx tz +
// g ty tz + h tx ty tz (8 float4 coefficients in registers), and runs every
// RK4 step of that z-cell from it: a stage costs 7 FMAs per channel and no
// memory access instead of 8 float4 gathers.  Each stage checks that its point
// lies in the cell; when one does not (the ray crosses an x/y face, or lies
// outside the grid in x/y), the step is redone with per-stage gathers and the
// cell is located again.  The cell is (re)loaded at the top of the outer loop
// only, so the coefficients are never merged across a branch.
// ---------------------------------------------------------------------------
struct Cell {
    float4 a, b, c, d, e, f, g, h;
};

__device__ __forceinline__ float4 f4sub(float4 p, float4 q) { return make_float4(p.x - q.x, p.y - q.y, p.z - q.z, p.w - q.w); }
__device__ __forceinline__ float4 f4fma(float4 p, float t, float4 q) {
    return make_float4(fmaf(p.x, t, q.x), fmaf(p.y, t, q.y), fmaf(p.z, t, q.z), fmaf(p.w, t, q.w));
}

__device__ __forceinline__ void load_cell(const TraceArgs &A, int ix, int iy, int iz, Cell &C) {
    const int64_t plane = (int64_t)A.nx * A.ny;
    const float4 *p = A.field + node_index(A, ix, iy, iz);
    const float4 c000 = p[0], c100 = p[1], c010 = p[A.nx], c110 = p[A.nx + 1];
    const float4 c001 = p[plane], c101 = p[plane + 1], c011 = p[plane + A.nx], c111 = p[plane + A.nx + 1];
    C.a = c000;
    C.b = f4sub(c100, c000);
    C.c = f4sub(c010, c000);
    C.d = f4sub(c001, c000);
    C.e = f4sub(f4sub(c110, c010), C.b);
    C.f = f4sub(f4sub(c101, c001), C.b);
    C.g = f4sub(f4sub(c011, c001), C.c);
    C.h = f4sub(f4sub(f4sub(c111, c011), f4sub(c101, c001)), C.e);
}

__device__ __forceinline__ float4 eval_cell(const Cell &C, float tx, float ty, float tz) {
    const float4 A = f4fma(C.d, tz, C.a), B = f4fma(C.f, tz, C.b);
    const float4 Cc = f4fma(C.g, tz, C.c), E = f4fma(C.h, tz, C.e);
    // A + tx (B + E ty) + Cc ty
    return f4fma(f4fma(E, ty, B), tx, f4fma(Cc, ty, A));
}

// one RK4 stage from the cached cell; ok &= the stage point lies in the cell
template <bool TAU, bool VF>
__device__ __forceinline__ void rhs_c(const TraceArgs &a, const Cell &C, float ox, float oy, float tz, const State &s,
                                      State &k, float &gw, float &vf, bool &ok) {
    const float tx = fmaf(a.sx, s.qx, ox), ty = fmaf(a.sy, s.qy, oy);  // ox = pxb - ix
    ok = ok && tx >= 0.f && tx <= 1.f && ty >= 0.f && ty <= 1.f;
    const float4 g = eval_cell(C, tx, ty, tz);
    const float W = 1.0f + s.Wd;
    const float iW = __builtin_amdgcn_rcpf(W);
    k.qx = s.U * iW;
    k.qy = s.V * iW;
    k.U = g.x * iW;
    k.V = g.y * iW;
    k.Wd = g.z * iW;
    k.tr = TAU ? -s.Wd * iW : 0.f;
    gw = g.w;
    if (VF) vf = path_factor(s, W, iW);
}

template <bool TAU>
__device__ __forceinline__ float rk4_step_c(const TraceArgs &a, const Cell &C, float ox, float oy, float tz0, float hh,
                                            State &y, bool &ok) {
    // running sum acc = k1 + 2 k2 + 2 k3 + k4 keeps one stage's slopes live
    State k, acc;
    float p, pacc, vf = 0.f;
    const float h2 = 0.5f * hh;
    rhs_c<TAU, false>(a, C, ox, oy, tz0, y, k, p, vf, ok);
    acc = k;
    pacc = p;
    rhs_c<TAU, false>(a, C, ox, oy, tz0 + h2, axpy(y, h2, k), k, p, vf, ok);
    acc = axpy(acc, 2.f, k);
    pacc = fmaf(2.f, p, pacc);
    rhs_c<TAU, true>(a, C, ox, oy, tz0 + h2, axpy(y, h2, k), k, p, vf, ok);
    acc = axpy(acc, 2.f, k);
    pacc = fmaf(2.f, p, pacc);
    rhs_c<TAU, false>(a, C, ox, oy, tz0 + hh, axpy(y, hh, k), k, p, vf, ok);
    acc = axpy(acc, 1.f, k);
    pacc += p;
    const float h6 = hh * (1.0f / 6.0f);
    y.qx = fmaf(h6, acc.qx, y.qx);
    y.qy = fmaf(h6, acc.qy, y.qy);
    y.U = fmaf(h6, acc.U, y.U);
    y.V = fmaf(h6, acc.V, y.V);
    y.Wd = fmaf(h6, acc.Wd, y.Wd);
    if (TAU) y.tr = fmaf(h6, acc.tr, y.tr);
    return h6 * pacc * vf;
}

// ---------------------------------------------------------------------------
// Two steps per cell (h = 1/2, the default).  The eight stages of a cell sit
// on five z-planes only (tz = 0, 1/4, 1/2, 3/4, 1), so the cell polynomial is
// reduced to each plane's bilinear form once (ZP: 4 FMAs per coefficient
// pair, none at tz = 0) and every stage evaluates that bilinear form (3 FMAs
// per channel).  Whether every stage point lies in the cell is tracked as the
// running min / max of tx, ty over the stages (two v_min3 / v_max3 per stage)
// and tested once per step.
// ---------------------------------------------------------------------------
struct ZP {
    float4 A, B, Cc, E;  // f = A + B tx + Cc ty + E tx ty on one z-plane
};

// The right-hand side at stage state s on plane z, unscaled: the derivatives
// are r / W with r = (U, V, gx, gy, gz, -Wd) (s's velocity and the field
// g), and 1 / W is returned in iW, so each use folds its RK4 weight into
// the one factor c / W (one multiply) instead of scaling the five
// derivatives first.  With SEED, iW comes in as the previous stage's 1 / W
// and is refined by one Newton step instead of a v_rcp (which costs several
// issue slots): its relative error e = 1 - W iW becomes e^2, within fp32
// rounding while |e| <= 2^-12; e is returned so that a step whose W changed
// more than that between stages is redone exactly.
template <bool VF, bool SEED = false>
__device__ __forceinline__ void rhs_z(const TraceArgs &a, const ZP &z, float ox, float oy, const State &s, float4 &g,
                                      float &iW, float &vf, float &mn, float &mx, float &e) {
    const float tx = fmaf(a.sx, s.qx, ox), ty = fmaf(a.sy, s.qy, oy);
    mn = fminf(mn, fminf(tx, ty));
    mx = fmaxf(mx, fmaxf(tx, ty));
    g = f4fma(f4fma(z.E, ty, z.B), tx, f4fma(z.Cc, ty, z.A));
    const float W = 1.0f + s.Wd;
    if (SEED) {
        e = fmaf(-W, iW, 1.0f);
        iW = fmaf(iW, e, iW);
    } else {
        iW = __builtin_amdgcn_rcpf(W);
    }
    if (VF) vf = path_factor(s, W, iW);
}

// the largest stage-to-stage relative change of W a cached step may have
// (the Newton-refined 1 / W is then exact to fp32 rounding)
constexpr float kNewtonTolInv = 4096.0f;

// y + c r, r the unscaled right-hand side of stage state s with field g
template <bool TAU>
__device__ __forceinline__ State axpy_r(const State &y, float c, const State &s, const float4 &g) {
    return State{fmaf(c, s.U, y.qx), fmaf(c, s.V, y.qy), fmaf(c, g.x, y.U),
                 fmaf(c, g.y, y.V),  fmaf(c, g.z, y.Wd), TAU ? fmaf(c, -s.Wd, y.tr) : y.tr};
}

// One RK4 step of h = 1/2 over the planes z0 (start), zm (middle), z1 (end).
// Returns the phase integrand Phi |v| / W at the step's start point: the
// phase is summed by the composite Simpson rule over the march's lattice
// points (march_m2), so only the first stage needs the Phi channel and the
// others' Phi (and the Phi channel of the zm and z1 planes) is never formed.
// mn / mx collect the stage points' in-cell coordinates, and mx also
// 4096 x the largest relative change of W between stages: the step is good
// from the cache only if mn >= 0 and mx <= 1.  iWc: in, 1 / W of a state near
// the start (the previous step's last stage, or exact); out, the last
// stage's 1 / W, the next step's seed.
template <bool TAU>
__device__ __forceinline__ float rk4_half_z(const TraceArgs &a, const ZP &z0, const ZP &zm, const ZP &z1, float ox,
                                            float oy, State &y, float &mn, float &mx, float &iWc) {
    // the new state is accumulated as y + h/6 (k1 + 2 k2 + 2 k3 + k4) stage
    // by stage, each weight folded into that stage's factor c / W
    constexpr float h6 = 0.5f / 6.0f, h3 = 0.5f / 3.0f;
    float4 g;
    float iW = iWc, vf = 0.f, unused, e1, e2, e3, e4;
    rhs_z<true, true>(a, z0, ox, oy, y, g, iW, vf, mn, mx, e1);
    const float phi = g.w * vf;
    State yn = axpy_r<TAU>(y, h6 * iW, y, g);
    const State s2 = axpy_r<TAU>(y, 0.25f * iW, y, g);
    rhs_z<false, true>(a, zm, ox, oy, s2, g, iW, unused, mn, mx, e2);
    yn = axpy_r<TAU>(yn, h3 * iW, s2, g);
    const State s3 = axpy_r<TAU>(y, 0.25f * iW, s2, g);
    rhs_z<false, true>(a, zm, ox, oy, s3, g, iW, unused, mn, mx, e3);
    yn = axpy_r<TAU>(yn, h3 * iW, s3, g);
    const State s4 = axpy_r<TAU>(y, 0.5f * iW, s3, g);
    rhs_z<false, true>(a, z1, ox, oy, s4, g, iW, unused, mn, mx, e4);
    y = axpy_r<TAU>(yn, h6 * iW, s4, g);
    iWc = iW;
    mx = fmaxf(mx, kNewtonTolInv * fmaxf(fmaxf(fabsf(e1), fabsf(e2)), fmaxf(fabsf(e3), fabsf(e4))));
    return phi;
}

// The cell as plane forms: z0 on its lower z-plane and D = (upper - lower),
// so the plane at tz is z0 + tz D (same coefficients as load_cell: A, B, Cc,
// E = a, b, c, e and D = d, f, g, h).  With FORMS the two planes' forms are
// loaded as written by tt_plane_forms (the same subtractions, so the same
// values bit for bit) and only D is formed here.
template <bool FORMS>
__device__ __forceinline__ void load_planes(const TraceArgs &A, int ix, int iy, int iz, ZP &z0, ZP &D) {
    if (FORMS) {
        const int64_t plane4 = 4 * (int64_t)A.nx * A.ny;
        const float4 *p = A.forms + 4 * node_index(A, ix, iy, iz);
        z0 = ZP{p[0], p[1], p[2], p[3]};
        const float4 a1 = p[plane4], b1 = p[plane4 + 1], c1 = p[plane4 + 2], e1 = p[plane4 + 3];
        D = ZP{f4sub(a1, z0.A), f4sub(b1, z0.B), f4sub(c1, z0.Cc), f4sub(e1, z0.E)};
        return;
    }
    const int64_t plane = (int64_t)A.nx * A.ny;
    const float4 *p = A.field + node_index(A, ix, iy, iz);
    const float4 c000 = p[0], c100 = p[1], c010 = p[A.nx], c110 = p[A.nx + 1];
    const float4 c001 = p[plane], c101 = p[plane + 1], c011 = p[plane + A.nx], c111 = p[plane + A.nx + 1];
    z0.A = c000;
    z0.B = f4sub(c100, c000);
    z0.Cc = f4sub(c010, c000);
    z0.E = f4sub(f4sub(c110, c010), z0.B);
    D.A = f4sub(c001, c000);
    D.B = f4sub(f4sub(c101, c001), z0.B);
    D.Cc = f4sub(f4sub(c011, c001), z0.Cc);
    D.E = f4sub(f4sub(f4sub(c111, c011), f4sub(c101, c001)), z0.E);
}

// the plane dt above plane z (D = the cell's upper-minus-lower plane form)
__device__ __forceinline__ ZP zstep(const ZP &z, float dt, const ZP &D) {
    return ZP{f4fma(D.A, dt, z.A), f4fma(D.B, dt, z.B), f4fma(D.Cc, dt, z.Cc), f4fma(D.E, dt, z.E)};
}

// Phase bookkeeping of the h = 1/2 march, per lane.  The phase
// int Phi |v| / W dZ is summed cell by cell: a cell [c, c + 1] whose two
// steps both ran from the cell cache (no stage point left the cell, so the
// integrand is the smooth cubic the trilinear field makes along the ray) takes
// Simpson's rule on the ray's points, (f(c) + 4 f(c + 1/2) + f(c + 1)) / 6; a
// step that crossed a face (gathered) takes the RK4 weights; the other step
// of such a cell takes Simpson's rule on its half cell, with its midpoint
// integrand gathered at the midpoint one quarter cell back along the ray.  A
// step needs the Phi channel at its first stage only; f(c + 1) is owed to the
// next step's start point (pend).
struct PhaseLedger {
    float pend = 0.f;  // weight owed to the next step's start-point integrand
    float phi0 = 0.f;  // the cell's first step's start-point integrand
    bool s1 = false;   // the cell's first step ran from the cache
};

// the cell's first step (even k): start-point integrand phi, smooth = ran
// from the cache, dph = its RK4 phase (used when it did not)
__device__ __forceinline__ void ledger_first(PhaseLedger &L, double &ph, float phi, bool smooth, float dph) {
    float add = L.pend * phi;
    if (!smooth) add += dph;
    L.pend = 0.f;
    L.s1 = smooth;
    L.phi0 = phi;
    ph += (double)add;
}

// the cell's second step (odd k) from state ys0 (its start) to ys1 (its end)
__device__ __forceinline__ void ledger_second(const TraceArgs &a, float pxb, float pyb, int k, PhaseLedger &L,
                                              double &ph, float phi, bool smooth, float dph, const State &ys0,
                                              const State &ys1) {
    constexpr float k6 = 1.0f / 6.0f, k12 = 1.0f / 12.0f;
    const float Zh = (float)k * 0.5f;  // the cell's half plane
    float add;
    if (L.s1 && smooth) {
        add = k6 * fmaf(4.0f, phi, L.phi0);
        L.pend = k6;
    } else if (L.s1) {  // the first half took the cache, this one crossed a face
        const float pm = phi_at(a, pxb, pyb, Zh - 0.25f, ys0, 0.25f);
        add = k12 * (L.phi0 + fmaf(4.0f, pm, phi)) + dph;
        L.pend = 0.f;
    } else if (smooth) {  // the first half crossed a face, this one took the cache
        const float pm = phi_at(a, pxb, pyb, Zh + 0.25f, ys1, 0.25f);
        add = k12 * fmaf(4.0f, pm, phi);
        L.pend = k12;
    } else {
        add = dph;
        L.pend = 0.f;
    }
    ph += (double)add;
}

// Steps k .. K-1 for h = 1/2 (two steps per cell).  A cell is held as its
// lower plane form and the upper-minus-lower difference (32 registers); the
// stage planes are z0 + D/4, z0 + D/2, then stepped by D/4 from there.  The
// common case runs as a tight loop over whole cells (both steps from the
// cache).  When a stage point of any lane leaves its cell the wave leaves
// that loop by one uniform branch, and the failing lanes redo that step with
// gathers.  If the redone step was the cell's first, the next iteration runs
// the second step from the cache: the lanes that redid it locate the cell
// they are now in, the others reload their own cell (cell and planes are
// formed the same way on both routes, and the phase ledger follows each
// lane's own steps, so what a lane computes never depends on the other lanes
// of its wave).  A start between z-planes (odd k) takes one gathered step
// first (its phase by the RK4 weights).  Returns false if the ray is handed
// over (the rule above kWdMin; the side faces are checked at each cell's start
// and at the end).
template <bool TAU, bool FORMS>
__device__ __forceinline__ bool march_m2(const TraceArgs &a, float pxb, float pyb, int k, int K, State &ys,
                                         double &ph) {
    bool good = true;
    const bool in0 = lateral_inside(a, pxb, pyb, ys);  // x / y inside the grid at the start
    bool crossed = false;                              // a side face crossed since
    if (k < K && (k & 1)) {
        ph += (double)rk4_step(a, pxb, pyb, (float)k * 0.5f, 0.5f, ys);
        good = ys.Wd > kWdMin;
        ++k;
    }
    if (k >= K) return good && lateral_inside(a, pxb, pyb, ys) == in0;
    PhaseLedger L;
    // 1 / W seeding each cached step's Newton refinement: the previous step's
    // last stage, or exact after a gathered step (per lane, so that a lane's
    // results never depend on which other lanes of its wave redid a step)
    float iWc = __builtin_amdgcn_rcpf(1.0f + ys.Wd);
    float ox = 0.f, oy = 0.f;  // entry point relative to the current cell
    bool inside = false;       // the cell lies in the grid (x, y)
    bool moved = false;        // this lane's last step was redone because it left its cell
    // locate the cell of the current position
    auto locate = [&](int &ix, int &iy) {
        const float px = fmaf(a.sx, ys.qx, pxb), py = fmaf(a.sy, ys.qy, pyb);
        inside = px >= 0.f && px <= a.fx1 && py >= 0.f && py <= a.fy1;
        crossed = crossed || inside != in0;
        ix = inside ? min((int)px, a.nx - 2) : 0;
        iy = inside ? min((int)py, a.ny - 2) : 0;
        ox = pxb - (float)ix;
        oy = pyb - (float)iy;
    };
    while (k < K) {
        State yn;
        float phi;
        bool ok;
        if (k & 1) {  // the second step of a cell whose first step was redone
            int ix, iy;
            if (moved) {
                locate(ix, iy);
            } else {  // this lane's own cell (|rounding of ox| << 1/2)
                ix = (int)rintf(pxb - ox);
                iy = (int)rintf(pyb - oy);
            }
            ZP z0, D;
            load_planes<FORMS>(a, ix, iy, k >> 1, z0, D);
            float mn = inside ? 0.f : -1.f, mx = 0.f;
            const ZP zh = zstep(z0, 0.5f, D);
            const ZP z3 = zstep(zh, 0.25f, D);
            const ZP z1 = zstep(z3, 0.25f, D);
            yn = ys;
            phi = rk4_half_z<TAU>(a, zh, z3, z1, ox, oy, yn, mn, mx, iWc);
            ok = mn >= 0.f && mx <= 1.f;
        } else {  // whole cells from the cache
            bool done = false;
            for (;;) {
                int ix, iy;
                // raised wave priority while the cell's loads are issued, so
                // they leave early and their latency hides behind other waves'
                // arithmetic (raising it for the gathered redo instead was slower)
                __builtin_amdgcn_s_setprio(1);
                locate(ix, iy);
                ZP z0, D;
                load_planes<FORMS>(a, ix, iy, k >> 1, z0, D);
                __builtin_amdgcn_s_setprio(0);
                float mn = inside ? 0.f : -1.f, mx = 0.f;
                const ZP zq = zstep(z0, 0.25f, D);
                const ZP zh = zstep(z0, 0.5f, D);
                yn = ys;
                phi = rk4_half_z<TAU>(a, z0, zq, zh, ox, oy, yn, mn, mx, iWc);
                ok = mn >= 0.f && mx <= 1.f;
                if (__ballot(!ok) != 0) break;  // redo the cell's first step below
                ys = yn;
                ledger_first(L, ph, phi, true, 0.f);
                ++k;
                good = good && (ys.Wd > kWdMin);
                const ZP z3 = zstep(zh, 0.25f, D);
                const ZP z1 = zstep(z3, 0.25f, D);
                yn = ys;
                phi = rk4_half_z<TAU>(a, zh, z3, z1, ox, oy, yn, mn, mx, iWc);
                ok = mn >= 0.f && mx <= 1.f;
                if (__ballot(!ok) != 0) break;  // redo the cell's second step below
                ledger_second(a, pxb, pyb, k, L, ph, phi, true, 0.f, ys, yn);
                ys = yn;
                ++k;
                good = good && (ys.Wd > kWdMin);
                if (k >= K) {
                    done = true;
                    break;
                }
            }
            if (done) break;
        }
        // step k: the lanes that left their cell redo it with gathers
        float dph = 0.f;
        if (!ok) {
            yn = ys;
            phi = rk4_step_phi(a, pxb, pyb, (float)k * 0.5f, 0.5f, yn, dph);
            iWc = __builtin_amdgcn_rcpf(1.0f + yn.Wd);  // an exact seed after a gathered step
        }
        if (k & 1)
            ledger_second(a, pxb, pyb, k, L, ph, phi, ok, dph, ys, yn);
        else
            ledger_first(L, ph, phi, ok, dph);
        moved = !ok;
        ys = yn;
        ++k;
        good = good && (ys.Wd > kWdMin);
    }
    if (L.pend != 0.f) ph += (double)(L.pend * phi_at(a, pxb, pyb, (float)K * 0.5f, ys));  // exit point
    return good && !crossed && lateral_inside(a, pxb, pyb, ys) == in0;
}

// Steps k .. K-1 (z = k h) from the lattice, cell-cached.  Every lane loads
// its cell (indices clamped to the grid, so a lane outside the grid or with a
// NaN position still reads valid memory) and runs the fast step; when a stage
// of any lane leaves its cell the wave takes one uniform branch in which the
// failing lanes redo the step with gathers, then every lane locates its cell
// again.  The cached coefficients are dead inside that branch, so the gather
// path does not add to the register footprint of the fast path.  Returns
// false if the ray is handed over (the rule above kWdMin).
template <bool TAU>
__device__ __forceinline__ bool march_cached(const TraceArgs &a, float pxb, float pyb, int k, int K, State &ys,
                                             double &ph) {
    bool good = true;
    const bool in0 = lateral_inside(a, pxb, pyb, ys);  // x / y inside the grid at the start
    while (k < K) {
        const int iz = k / a.m;
        const int kend = min(K, (iz + 1) * a.m);
        const float px = fmaf(a.sx, ys.qx, pxb), py = fmaf(a.sy, ys.qy, pyb);
        const bool inside = px >= 0.f && px <= a.fx1 && py >= 0.f && py <= a.fy1;
        good = good && inside == in0;
        const int ix = inside ? min((int)px, a.nx - 2) : 0;
        const int iy = inside ? min((int)py, a.ny - 2) : 0;
        Cell C;
        load_cell(a, ix, iy, iz, C);
        const float ox = pxb - (float)ix, oy = pyb - (float)iy;
        bool redo = false, ok = true;
        State yn;
        float dph = 0.f;
        while (!redo && k < kend) {
            const float tz0 = (float)(k - iz * a.m) * a.h;
            ok = inside;
            yn = ys;
            dph = rk4_step_c<TAU>(a, C, ox, oy, tz0, a.h, yn, ok);
            if (__ballot(!ok) != 0) {  // wave-uniform: redo step k below
                redo = true;
                break;
            }
            ys = yn;
            ph += (double)dph;
            ++k;
            good = good && (ys.Wd > kWdMin);
        }
        if (redo) {  // the cell is dead here; failing lanes gather
            if (!ok) {
                yn = ys;
                dph = rk4_step(a, pxb, pyb, (float)k * a.h, a.h, yn);
            }
            ys = yn;
            ph += (double)dph;
            ++k;
            good = good && (ys.Wd > kWdMin);
        }
    }
    return good && lateral_inside(a, pxb, pyb, ys) == in0;
}

// One ray: the t-th of the launch (a.order[t] when given).  CACHED selects
// the cell-cached march (M2: its h = 1/2 form, which requires a.m == 2;
// FORMS: its planes from the plane-form field a.forms).
template <bool CACHED, bool M2, bool TAU, bool FORMS = false>
__device__ __forceinline__ void trace_ray(const TraceArgs &a, int64_t t) {
    const int64_t r = a.order ? (int64_t)a.order[t] : t;
    const double *s = a.s0 + r;
    const double x = s[0], y = s[a.ld_s0], z = s[2 * a.ld_s0];
    const double vx = s[3 * a.ld_s0], vy = s[4 * a.ld_s0], vz = s[5 * a.ld_s0];
    int st = 0;
    if (isnan(x) || isnan(y) || isnan(z) || isnan(vx) || isnan(vy) || isnan(vz))
        st = 2;
    else if (!(vz / tt::kC - 1.0 > (double)kWdMin))  // handed over at launch
        st = 1;

    const double U0 = vx / tt::kC, V0 = vy / tt::kC, W0 = vz / tt::kC;
    double px = (x - a.x0) / a.dx, py = (y - a.y0) / a.dy, Zs = (z - a.z0) / a.dz;
    double tau = 0.0;
    const double Zend = (double)(a.nz - 1);
    State ys{0.f, 0.f, (float)U0, (float)V0, (float)(W0 - 1.0), 0.f};
    double ph = 0.0;

    if (st == 0) {
        // field is zero before the first plane: carry the ray straight there
        if (Zs < 0.0) {
            const double d = -Zs;
            px += (double)a.sx * U0 / W0 * d;
            py += (double)a.sy * V0 / W0 * d;
            tau += a.dz * d / W0;
            Zs = 0.0;
        }
        if (Zs >= Zend) {
            // starts at or past the exit plane: straight back-projection
            const double d = Zend - Zs;
            px += (double)a.sx * U0 / W0 * d;
            py += (double)a.sy * V0 / W0 * d;
            tau += a.dz * d / W0;
        } else {
            const float pxb = (float)px, pyb = (float)py;
            const int K = (a.nz - 1) * a.m;
            const int kstart = (int)ceil(Zs * a.m - 1e-9);
            const double h0 = kstart * (double)a.h - Zs;
            if (CACHED) {
                // partial first step (off-lattice start) with gathers, then the lattice
                if (h0 > 1e-12) ph += (double)rk4_step(a, pxb, pyb, (float)Zs, (float)h0, ys);
                if (!(M2 ? march_m2<TAU, FORMS>(a, pxb, pyb, kstart, K, ys, ph)
                         : march_cached<TAU>(a, pxb, pyb, kstart, K, ys, ph)))
                    st = 1;
            } else {
                if (h0 > 1e-12) ph += (double)rk4_step(a, pxb, pyb, (float)Zs, (float)h0, ys);
                const bool in0 = lateral_inside(a, pxb, pyb, ys);
                for (int k = kstart; k < K; ++k) {
                    ph += (double)rk4_step(a, pxb, pyb, (float)k * a.h, a.h, ys);
                    if (!(ys.Wd > kWdMin) || lateral_inside(a, pxb, pyb, ys) != in0) {  // handed over
                        st = 1;
                        break;
                    }
                }
            }
            px += (double)a.sx * (double)ys.qx;
            py += (double)a.sy * (double)ys.qy;
            tau += a.dz * ((Zend - Zs) + (double)ys.tr);
        }
    }

    const double nanv = __longlong_as_double(0x7ff8000000000000ll);
    const double U = ys.U, V = ys.V, W = 1.0 + (double)ys.Wd;
    const double xe = a.x0 + px * a.dx, ye = a.y0 + py * a.dy, ze = a.z0 + Zend * a.dz;
    const bool ok = st == 0;
    double *rf = a.rf + r;
    rf[0] = ok ? xe : nanv;
    rf[a.ld_rf] = ok ? atan(U / W) : nanv;
    rf[2 * a.ld_rf] = ok ? ye : nanv;
    rf[3 * a.ld_rf] = ok ? atan(V / W) : nanv;
    if (a.phase) a.phase[r] = ok ? ph : nanv;
    if (a.sf) {
        const double rem = a.cT - tau;
        double *sf = a.sf + r;
        sf[0] = ok ? xe + U * rem : nanv;
        sf[a.ld_sf] = ok ? ye + V * rem : nanv;
        sf[2 * a.ld_sf] = ok ? ze + W * rem : nanv;
        sf[3 * a.ld_sf] = ok ? tt::kC * U : nanv;
        sf[4 * a.ld_sf] = ok ? tt::kC * V : nanv;
        sf[5 * a.ld_sf] = ok ? tt::kC * W : nanv;
    }
    if (a.status) a.status[r] = st;
}

// The h = 1/2 form fits 96 VGPRs (five waves per SIMD) with one 8-byte spill
// outside the step loop.
template <bool CACHED, bool M2, bool TAU, bool FORMS = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(M2 ? 5 : 1))) void trace_rk4_kernel(TraceArgs a) {
    const int b = tt::xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int64_t t = (int64_t)b * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    if (a.n_dev && t >= (int64_t)*a.n_dev) return;
    trace_ray<CACHED, M2, TAU, FORMS>(a, t);
}

// ------------------------------------------------------------- ensembles
// Many realisations in one persistent launch.  The rays of every item are
// cut into chunks (ordered longest rays first); each block takes the next
// chunk from a device counter until none is left, so uneven grids and ray
// counts balance over the whole chip without a launch per item.
struct BatchHeader {
    unsigned long long counter;  // next chunk; reset by reset_kernel before each launch
};

__global__ void reset_kernel(BatchHeader *h) {
    if (threadIdx.x == 0 && blockIdx.x == 0) h->counter = 0ull;
}

template <bool CACHED, bool M2>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(M2 ? 5 : 4))) void trace_batch_kernel(BatchHeader *hdr, const TraceArgs *__restrict__ items,
                                                          const int32_t *__restrict__ chunk_item,
                                                          const int64_t *__restrict__ chunk_start, int nchunks,
                                                          int chunk) {
    __shared__ int next;
    for (;;) {
        __syncthreads();  // every lane has read `next` of the previous round
        if (threadIdx.x == 0) next = (int)atomicAdd(&hdr->counter, 1ull);
        __syncthreads();
        const int c = __builtin_amdgcn_readfirstlane(next);
        if (c >= nchunks) return;  // block-uniform: every wave of the block leaves
        const TraceArgs &a = items[__builtin_amdgcn_readfirstlane(chunk_item[c])];
        const int64_t s = chunk_start[c], e = min(s + (int64_t)chunk, a.n);
        for (int64_t t = s + threadIdx.x; t < e; t += blockDim.x) trace_ray<CACHED, M2, false>(a, t);
    }
}

// ------------------------------------------------------------ rays in time
// Rays the z-march cannot follow (launched with vz <= 0, or turned back so
// that W <= 0 on the way, status 1) are integrated as the reference does, in
// time (PT:312-331, solve_ivp's RK45), but per ray and at a tight tolerance:
// the Dormand-Prince 5(4) embedded pair (Dormand & Prince, J. Comput. Appl.
// Math. 6 (1980) 19-26; the pair scipy's RK45 uses) in tau = c t on
//   dX/dtau = U,  dU/dtau = g / dz,  dphase/dtau = (Phi / dz) |U|
// (X in metres, U = v / c, g and Phi the node field's channels, which carry
// one dz) from tau = 0 to cT.  A step is accepted when every component's
// error estimate is within kTimeTol of its scale (dx, dy, dz for positions,
// 1 for U, max(1, |phase|) for the phase); steps are at most half the
// smallest cell edge and shrink where the trilinear field's kinks at cell
// faces need it.  The state and the trilinear weights are fp64: these are few
// rays, near a turning point W is small and its relative precision matters,
// and fp32 cell coordinates would put noise of 1e-7 x n cells into the field
// that a tolerance this tight would chase.  Once a ray is outside the grid
// and moving away from it on that axis the field can no longer act, and it is
// carried straight to cT.  A ray that needs more than kTimeMaxSteps attempted
// steps is given up (NaN outputs, status 2).
constexpr double kTimeTol = 1e-10;
constexpr int kTimeMaxSteps = 1 << 18;

__device__ __forceinline__ void time_rhs(const TraceArgs &a, const double (&y)[7], double (&k)[7]) {
    const double px = (y[0] - a.x0) / a.dx, py = (y[1] - a.y0) / a.dy, pz = (y[2] - a.z0) / a.dz;
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    if (px >= 0.0 && px <= (double)a.fx1 && py >= 0.0 && py <= (double)a.fy1 && pz >= 0.0 && pz <= (double)a.fz1) {
        const int ix = min((int)px, a.nx - 2), iy = min((int)py, a.ny - 2), iz = min((int)pz, a.nz - 2);
        const double tx = px - ix, ty = py - iy, tz = pz - iz;
        const int64_t plane = (int64_t)a.nx * a.ny;
        const float4 *p = a.field + node_index(a, ix, iy, iz);
        const int64_t off[8] = {0, 1, a.nx, a.nx + 1, plane, plane + 1, plane + a.nx, plane + a.nx + 1};
        for (int c = 0; c < 8; ++c) {
            const double w = ((c & 1) ? tx : 1.0 - tx) * ((c & 2) ? ty : 1.0 - ty) * ((c & 4) ? tz : 1.0 - tz);
            const float4 v = p[off[c]];
            g[0] = fma(w, (double)v.x, g[0]);
            g[1] = fma(w, (double)v.y, g[1]);
            g[2] = fma(w, (double)v.z, g[2]);
            g[3] = fma(w, (double)v.w, g[3]);
        }
    }
    const double idz = 1.0 / a.dz;
    k[0] = y[3];
    k[1] = y[4];
    k[2] = y[5];
    k[3] = g[0] * idz;
    k[4] = g[1] * idz;
    k[5] = g[2] * idz;
    k[6] = g[3] * idz * sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
}

// Dormand-Prince 5(4) tableau: stage nodes are implied by the rows; the last
// row is the fifth-order solution (and its stage the next step's first, FSAL)
__constant__ double kDPA[6][6] = {
    {1.0 / 5, 0, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0},
    {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84}};
// fifth-order minus fourth-order weights, over the seven stages
__constant__ double kDPE[7] = {71.0 / 57600, 0, -71.0 / 16695, 71.0 / 1920, -17253.0 / 339200, 22.0 / 525, -1.0 / 40};

// ray t of a (the order's t-th) if its status is `only` (or a has none)
__device__ void trace_time_ray(const TraceArgs &a, int64_t t, int32_t only) {
    const int64_t r = a.order ? (int64_t)a.order[t] : t;
    if (a.status && a.status[r] != only) return;
    const double *s = a.s0 + r;
    double y[7] = {s[0], s[a.ld_s0], s[2 * a.ld_s0], s[3 * a.ld_s0] / tt::kC, s[4 * a.ld_s0] / tt::kC,
                   s[5 * a.ld_s0] / tt::kC, 0.0};
    const double nanv = __longlong_as_double(0x7ff8000000000000ll);
    bool bad = false;
    for (int i = 0; i < 6; ++i) bad = bad || isnan(y[i]);
    if (!bad) {
        const double hmax = 0.5 * fmin(a.dx, fmin(a.dy, a.dz));
        const double lo[3] = {a.x0, a.y0, a.z0};
        const double hi[3] = {a.x0 + (double)a.fx1 * a.dx, a.y0 + (double)a.fy1 * a.dy, a.z0 + (double)a.fz1 * a.dz};
        const double sc[6] = {kTimeTol * a.dx, kTimeTol * a.dy, kTimeTol * a.dz, kTimeTol, kTimeTol, kTimeTol};
        double k[7][7];
        time_rhs(a, y, k[0]);
        double tau = 0.0, h = hmax;
        int tries = 0;
        while (tau < a.cT) {
            bool gone = false;  // outside the grid and moving away from it
            for (int d = 0; d < 3; ++d)
                gone = gone || (y[d] < lo[d] && y[3 + d] <= 0.0) || (y[d] > hi[d] && y[3 + d] >= 0.0);
            if (gone) break;
            if (++tries > kTimeMaxSteps) {
                bad = true;
                break;
            }
            const double hh = fmin(h, a.cT - tau);
            double yn[7];
#pragma unroll
            for (int st = 1; st <= 6; ++st) {
#pragma unroll
                for (int i = 0; i < 7; ++i) {
                    double acc = 0.0;
#pragma unroll
                    for (int j = 0; j < st; ++j) acc = fma(kDPA[st - 1][j], k[j][i], acc);
                    yn[i] = fma(hh, acc, y[i]);
                }
                time_rhs(a, yn, k[st]);
            }
            double en = 0.0;  // largest scaled error estimate
#pragma unroll
            for (int i = 0; i < 7; ++i) {
                double e = 0.0;
#pragma unroll
                for (int j = 0; j < 7; ++j) e = fma(kDPE[j], k[j][i], e);
                const double scale = i < 6 ? sc[i] : kTimeTol * fmax(1.0, fmax(fabs(y[6]), fabs(yn[6])));
                en = fmax(en, fabs(hh * e) / scale);
            }
            double fac = 0.9 * pow(fmax(en, 1e-10), -0.2);
            fac = fmin(5.0, fmax(0.2, fac));
            if (en <= 1.0) {  // accept: the last stage is the next step's first
#pragma unroll
                for (int i = 0; i < 7; ++i) {
                    y[i] = yn[i];
                    k[0][i] = k[6][i];
                }
                tau += hh;
            } else {
                fac = fmin(fac, 1.0);
            }
            h = fmin(hh * fac, hmax);
        }
        if (!bad && tau < a.cT)  // straight to cT (no field acts any more)
            for (int d = 0; d < 3; ++d) y[d] = fma(a.cT - tau, y[3 + d], y[d]);
    }
    // ray_at_exit (PT:333-380): back-project to the exit plane z = z.max()
    const double ze = a.z0 + (double)a.fz1 * a.dz;
    const double tbp = (y[2] - ze) / y[5];
    double *rf = a.rf + r;
    rf[0] = bad ? nanv : y[0] - y[3] * tbp;
    rf[a.ld_rf] = bad ? nanv : atan(y[3] / y[5]);
    rf[2 * a.ld_rf] = bad ? nanv : y[1] - y[4] * tbp;
    rf[3 * a.ld_rf] = bad ? nanv : atan(y[4] / y[5]);
    if (a.phase) a.phase[r] = bad ? nanv : y[6];
    if (a.sf) {
        double *sf = a.sf + r;
        for (int i = 0; i < 6; ++i) sf[i * a.ld_sf] = bad ? nanv : (i < 3 ? y[i] : tt::kC * y[i]);
    }
    if (a.status) a.status[r] = bad ? 2 : 3;
}

__global__ __launch_bounds__(256) void trace_time_kernel(TraceArgs a, int32_t only) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.n) trace_time_ray(a, t, only);
}

// the rays of a batch the z-march handed over (status 1): one block per chunk
// of the batch's chunk table, items without a status array skipped
__global__ __launch_bounds__(256) void trace_time_batch_kernel(const TraceArgs *__restrict__ items,
                                                               const int32_t *__restrict__ chunk_item,
                                                               const int64_t *__restrict__ chunk_start, int chunk) {
    const TraceArgs &a = items[chunk_item[blockIdx.x]];
    if (!a.status) return;
    const int64_t s = chunk_start[blockIdx.x], e = min(s + (int64_t)chunk, a.n);
    for (int64_t t = s + threadIdx.x; t < e; t += blockDim.x) trace_time_ray(a, t, 1);
}

// trilinear samples of the node field at arbitrary points (metres)
__global__ __launch_bounds__(256) void sample_kernel(TraceArgs a, const double *pts, int64_t ld, int64_t n,
                                                     double *out, int64_t ldo) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pts[i], y = pts[ld + i], z = pts[2 * ld + i];
    const double nanv = __longlong_as_double(0x7ff8000000000000ll);
    if (isnan(x) || isnan(y) || isnan(z)) {
        for (int c = 0; c < 4; ++c) out[c * ldo + i] = nanv;
        return;
    }
    // bounds in double (RegularGridInterpolator compares coordinates)
    const double px = (x - a.x0) / a.dx, py = (y - a.y0) / a.dy, pz = (z - a.z0) / a.dz;
    const bool in = px >= 0.0 && px <= (double)a.fx1 && py >= 0.0 && py <= (double)a.fy1 && pz >= 0.0 &&
                    pz <= (double)a.fz1;
    float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
    if (in) g = sample(a, fminf((float)px, a.fx1), fminf((float)py, a.fy1), fminf((float)pz, a.fz1));
    out[i] = g.x;
    out[ldo + i] = g.y;
    out[2 * ldo + i] = g.z;
    out[3 * ldo + i] = g.w;
}

int fill_grid_args(const tt_grid *grid, TraceArgs &a) {
    TT_REQUIRE(grid->nx >= 2 && grid->ny >= 2 && grid->nz >= 2, "grid needs >= 2 nodes per axis");
    TT_REQUIRE(grid->dx > 0 && grid->dy > 0 && grid->dz > 0, "grid spacing must be > 0");
    TT_REQUIRE((int64_t)grid->nx * grid->ny < ((int64_t)1 << 31), "grid: nx * ny must be < 2^31, got %d x %d",
               grid->nx, grid->ny);
    a.nx = grid->nx;
    a.ny = grid->ny;
    a.nz = grid->nz;
    a.plane = (uint32_t)grid->nx * (uint32_t)grid->ny;
    a.fx1 = (float)(grid->nx - 1);
    a.fy1 = (float)(grid->ny - 1);
    a.fz1 = (float)(grid->nz - 1);
    a.x0 = grid->x0;
    a.y0 = grid->y0;
    a.z0 = grid->z0;
    a.dx = grid->dx;
    a.dy = grid->dy;
    a.dz = grid->dz;
    a.sx = (float)(grid->dz / grid->dx);
    a.sy = (float)(grid->dz / grid->dy);
    return TT_OK;
}

}  // namespace

extern "C" int tt_sample_field(const float *field4, const tt_grid *grid, const double *pts, int64_t ld, int64_t n,
                               double *out, int64_t ld_out, void *stream) {
    TT_REQUIRE(grid, "tt_sample_field: null grid");
    TT_REQUIRE(n >= 0 && ld >= n && ld_out >= n, "tt_sample_field: bad sizes");
    if (n == 0) return TT_OK;  // an empty point set: nothing is read or written
    TT_REQUIRE(field4 && pts && out, "tt_sample_field: null pointer");
    TraceArgs a{};
    if (int rc = fill_grid_args(grid, a)) return rc;
    a.field = reinterpret_cast<const float4 *>(field4);
    hipLaunchKernelGGL(sample_kernel, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0, tt::as_stream(stream), a,
                       pts, ld, n, out, ld_out);
    TT_LAUNCH_CHECK("tt_sample_field launch");
    return TT_OK;
}

namespace {
int trace_launch(const float *field4, const tt_grid *grid, const tt_trace_opts *opts, const double *s0, int64_t ld_s0,
                 int64_t n, const int32_t *order, const int32_t *n_dev, double *rf, int64_t ld_rf, double *phase,
                 double *sf, int64_t ld_sf, int32_t *status, void *stream) {
    TT_REQUIRE(grid && opts, "tt_trace_rk4: null grid or options");
    TT_REQUIRE(n >= 0, "tt_trace_rk4: n < 0");
    if (n == 0) return TT_OK;  // an empty ray set: nothing is read or written
    TT_REQUIRE(field4 && s0 && rf, "tt_trace_rk4: null pointer");
    TT_REQUIRE(ld_s0 >= n && ld_rf >= n && (!sf || ld_sf >= n), "tt_trace_rk4: leading dimension < n");
    const double mm = 1.0 / opts->h;
    const int m = (int)llround(mm);
    TT_REQUIRE(m >= 1 && m <= 64 && fabs(mm - m) < 1e-9, "tt_trace_rk4: h must be 1/m cells with 1 <= m <= 64");
    TT_REQUIRE((int64_t)(grid->nz - 1) * m < (1 << 23), "tt_trace_rk4: too many steps for exact float Z");
    TraceArgs a{};
    if (int rc = fill_grid_args(grid, a)) return rc;
    a.field = reinterpret_cast<const float4 *>(field4);
    a.m = m;
    a.h = 1.0f / (float)m;
    a.cT = opts->cT;
    a.s0 = s0;
    a.ld_s0 = ld_s0;
    a.n = n;
    a.order = order;
    a.rf = rf;
    a.ld_rf = ld_rf;
    a.phase = phase;
    a.sf = sf;
    a.ld_sf = ld_sf;
    a.status = status;
    a.n_dev = n_dev;
    a.forms = reinterpret_cast<const float4 *>(opts->forms);
    TT_REQUIRE(!opts->forms || ((uintptr_t)opts->forms & 15) == 0, "tt_trace_rk4: forms must be 16-byte aligned");
    const int block = 256;
    TT_REQUIRE(opts->variant == 0 || opts->variant == 1, "tt_trace_rk4: variant must be 0 or 1");
    const dim3 nblk((unsigned)tt::ceil_div(n, block));
    hipStream_t st = tt::as_stream(stream);
    // the time integral is needed only for sf
    const bool m2 = m == 2;
    if (opts->variant == 1)
        hipLaunchKernelGGL((trace_rk4_kernel<false, false, true>), nblk, dim3(block), 0, st, a);
    else if (m2 && sf)
        hipLaunchKernelGGL((trace_rk4_kernel<true, true, true>), nblk, dim3(block), 0, st, a);
    else if (m2 && a.forms)
        hipLaunchKernelGGL((trace_rk4_kernel<true, true, false, true>), nblk, dim3(block), 0, st, a);
    else if (m2)
        hipLaunchKernelGGL((trace_rk4_kernel<true, true, false>), nblk, dim3(block), 0, st, a);
    else if (sf)
        hipLaunchKernelGGL((trace_rk4_kernel<true, false, true>), nblk, dim3(block), 0, st, a);
    else
        hipLaunchKernelGGL((trace_rk4_kernel<true, false, false>), nblk, dim3(block), 0, st, a);
    TT_LAUNCH_CHECK("tt_trace_rk4 launch");
    return TT_OK;
}
}  // namespace

extern "C" int tt_trace_rk4(const float *field4, const tt_grid *grid, const tt_trace_opts *opts, const double *s0,
                            int64_t ld_s0, int64_t n, const int32_t *order, double *rf, int64_t ld_rf,
                            double *phase, double *sf, int64_t ld_sf, int32_t *status, void *stream) {
    return trace_launch(field4, grid, opts, s0, ld_s0, n, order, nullptr, rf, ld_rf, phase, sf, ld_sf, status,
                        stream);
}

extern "C" int tt_trace_time(const float *field4, const tt_grid *grid, const tt_trace_opts *opts, const double *s0,
                             int64_t ld_s0, int64_t n, const int32_t *order, double *rf, int64_t ld_rf, double *phase,
                             double *sf, int64_t ld_sf, int32_t *status, int32_t only_status, void *stream) {
    TT_REQUIRE(grid && opts, "tt_trace_time: null grid or options");
    TT_REQUIRE(n >= 0, "tt_trace_time: n < 0");
    if (n == 0) return TT_OK;  // an empty ray set: nothing is read or written
    TT_REQUIRE(field4 && s0 && rf, "tt_trace_time: null pointer");
    TT_REQUIRE(ld_s0 >= n && ld_rf >= n && (!sf || ld_sf >= n), "tt_trace_time: leading dimension < n");
    TT_REQUIRE(opts->cT > 0.0, "tt_trace_time: cT must be > 0");
    TraceArgs a{};
    if (int rc = fill_grid_args(grid, a)) return rc;
    a.field = reinterpret_cast<const float4 *>(field4);
    a.cT = opts->cT;
    a.s0 = s0;
    a.ld_s0 = ld_s0;
    a.n = n;
    a.order = order;
    a.rf = rf;
    a.ld_rf = ld_rf;
    a.phase = phase;
    a.sf = sf;
    a.ld_sf = ld_sf;
    a.status = status;
    hipLaunchKernelGGL(trace_time_kernel, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0, tt::as_stream(stream), a,
                       only_status);
    TT_LAUNCH_CHECK("tt_trace_time launch");
    return TT_OK;
}

extern "C" int tt_trace_rk4_counted(const float *field4, const tt_grid *grid, const tt_trace_opts *opts,
                                    const double *s0, int64_t ld_s0, int64_t n_cap, const int32_t *order,
                                    const int32_t *n_dev, double *rf, int64_t ld_rf, double *phase, int32_t *status,
                                    void *stream) {
    TT_REQUIRE(n_cap == 0 || (order && n_dev), "tt_trace_rk4_counted: order and n_dev are required");
    return trace_launch(field4, grid, opts, s0, ld_s0, n_cap, order, n_dev, rf, ld_rf, phase, nullptr, ld_rf, status,
                        stream);
}

namespace {
constexpr size_t kBatchAlign = 256;
size_t align_up(size_t v) { return (v + kBatchAlign - 1) / kBatchAlign * kBatchAlign; }

struct BatchLayout {
    size_t items, citem, cstart, total;
};

BatchLayout batch_layout(int32_t nitems, int64_t nchunks) {
    BatchLayout L;
    L.items = align_up(sizeof(BatchHeader));
    L.citem = L.items + align_up(sizeof(TraceArgs) * (size_t)nitems);
    L.cstart = L.citem + align_up(sizeof(int32_t) * (size_t)nchunks);
    L.total = L.cstart + align_up(sizeof(int64_t) * (size_t)nchunks);
    return L;
}

int64_t count_chunks(const tt_batch_item *items, int32_t nitems, int32_t chunk) {
    int64_t c = 0;
    for (int i = 0; i < nitems; ++i) c += tt::ceil_div(items[i].n, (int64_t)chunk);
    return c;
}
}  // namespace

extern "C" int tt_batch_workspace_bytes(const tt_batch_item *items, int32_t nitems, int32_t chunk, size_t *bytes) {
    TT_REQUIRE(items && bytes && nitems >= 1 && chunk >= 1, "tt_batch_workspace_bytes: bad arguments");
    *bytes = batch_layout(nitems, count_chunks(items, nitems, chunk)).total;
    return TT_OK;
}

extern "C" int tt_batch_prepare(const tt_batch_item *items, int32_t nitems, double h, int32_t chunk, void *workspace,
                                size_t bytes, tt_batch_info *info) {
    TT_REQUIRE(items && workspace && info && nitems >= 1, "tt_batch_prepare: bad arguments");
    TT_REQUIRE(chunk >= 64 && chunk <= (1 << 24), "tt_batch_prepare: chunk must be in [64, 2^24]");
    const double mm = 1.0 / h;
    const int m = (int)llround(mm);
    TT_REQUIRE(m >= 1 && m <= 64 && fabs(mm - m) < 1e-9, "tt_batch_prepare: h must be 1/m cells with 1 <= m <= 64");
    const int64_t nchunks = count_chunks(items, nitems, chunk);
    TT_REQUIRE(nchunks >= 1 && nchunks < (1ll << 31), "tt_batch_prepare: no rays or too many chunks");
    const BatchLayout L = batch_layout(nitems, nchunks);
    TT_REQUIRE(bytes >= L.total, "tt_batch_prepare: workspace of %zu bytes < %zu needed", bytes, L.total);
    std::vector<TraceArgs> args((size_t)nitems);
    for (int i = 0; i < nitems; ++i) {
        const tt_batch_item &it = items[i];
        TT_REQUIRE(it.field4 && it.s0 && it.rf && it.n >= 0, "tt_batch_prepare: item %d: null pointer or n < 0", i);
        TT_REQUIRE(it.ld_s0 >= it.n && it.ld_rf >= it.n, "tt_batch_prepare: item %d: leading dimension < n", i);
        TT_REQUIRE((int64_t)(it.grid.nz - 1) * m < (1 << 23), "tt_batch_prepare: item %d: too many steps", i);
        TT_REQUIRE(!it.status || it.cT > 0.0, "tt_batch_prepare: item %d: cT must be > 0 with a status array", i);
        TraceArgs a{};
        if (int rc = fill_grid_args(&it.grid, a)) return rc;
        a.field = reinterpret_cast<const float4 *>(it.field4);
        a.m = m;
        a.h = 1.0f / (float)m;
        a.cT = it.status ? it.cT : 0.0;
        a.s0 = it.s0;
        a.ld_s0 = it.ld_s0;
        a.n = it.n;
        a.order = it.order;
        a.rf = it.rf;
        a.ld_rf = it.ld_rf;
        a.phase = it.phase;
        a.sf = nullptr;
        a.ld_sf = 0;
        a.status = it.status;
        args[(size_t)i] = a;
    }
    // chunks of the items with the longest rays (most z-steps) first
    std::vector<int32_t> idx((size_t)nitems);
    for (int i = 0; i < nitems; ++i) idx[(size_t)i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t p, int32_t q) { return items[p].grid.nz > items[q].grid.nz; });
    std::vector<int32_t> citem;
    std::vector<int64_t> cstart;
    citem.reserve((size_t)nchunks);
    cstart.reserve((size_t)nchunks);
    for (int32_t i : idx)
        for (int64_t s = 0; s < items[i].n; s += chunk) {
            citem.push_back(i);
            cstart.push_back(s);
        }
    BatchHeader hdr{0ull};
    char *w = static_cast<char *>(workspace);
    TT_HIP(hipMemcpy(w, &hdr, sizeof(hdr), hipMemcpyHostToDevice), "tt_batch_prepare: header upload");
    TT_HIP(hipMemcpy(w + L.items, args.data(), sizeof(TraceArgs) * args.size(), hipMemcpyHostToDevice),
           "tt_batch_prepare: item upload");
    TT_HIP(hipMemcpy(w + L.citem, citem.data(), sizeof(int32_t) * citem.size(), hipMemcpyHostToDevice),
           "tt_batch_prepare: chunk table upload");
    TT_HIP(hipMemcpy(w + L.cstart, cstart.data(), sizeof(int64_t) * cstart.size(), hipMemcpyHostToDevice),
           "tt_batch_prepare: chunk table upload");
    info->nitems = nitems;
    info->nchunks = (int32_t)nchunks;
    info->chunk = chunk;
    info->steps_per_cell = m;
    return TT_OK;
}

extern "C" int tt_batch_trace(void *workspace, const tt_batch_info *info, const tt_batch_tuning *tune, void *stream) {
    TT_REQUIRE(workspace && info && tune, "tt_batch_trace: null pointer");
    TT_REQUIRE(info->nitems >= 1 && info->nchunks >= 1 && info->chunk >= 64, "tt_batch_trace: info not prepared");
    TT_REQUIRE(tune->block == 64 || tune->block == 128 || tune->block == 256,
               "tt_batch_trace: block must be 64, 128 or 256");
    TT_REQUIRE(tune->variant == 0 || tune->variant == 1, "tt_batch_trace: variant must be 0 or 1");
    TT_REQUIRE(tune->blocks_per_cu >= 0 && tune->blocks_per_cu <= 64, "tt_batch_trace: blocks_per_cu in [0, 64]");
    TT_REQUIRE(info->steps_per_cell >= 1 && info->steps_per_cell <= 64, "tt_batch_trace: info not prepared");
    // variant 1: gathers; variant 0: cell-cached (its h = 1/2 form when m == 2)
    using BatchKernel = void (*)(BatchHeader *, const TraceArgs *, const int32_t *, const int64_t *, int, int);
    const BatchKernel kern = tune->variant == 1          ? trace_batch_kernel<false, false>
                             : info->steps_per_cell == 2 ? trace_batch_kernel<true, true>
                                                         : trace_batch_kernel<true, false>;
    int dev = 0, ncu = 0, occ = 0;
    TT_HIP(hipGetDevice(&dev), "tt_batch_trace: hipGetDevice");
    TT_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "tt_batch_trace: CU count");
    TT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, tune->block, 0), "tt_batch_trace: occupancy");
    const int per_cu = tune->blocks_per_cu > 0 ? std::min(tune->blocks_per_cu, std::max(occ, 1)) : std::max(occ, 1);
    const dim3 grid((unsigned)(ncu * per_cu));
    char *w = static_cast<char *>(workspace);
    BatchHeader *hdr = reinterpret_cast<BatchHeader *>(w);
    const BatchLayout L = batch_layout(info->nitems, info->nchunks);
    hipStream_t st = tt::as_stream(stream);
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, st, hdr);
    const TraceArgs *items = reinterpret_cast<const TraceArgs *>(w + L.items);
    const int32_t *citem = reinterpret_cast<const int32_t *>(w + L.citem);
    const int64_t *cstart = reinterpret_cast<const int64_t *>(w + L.cstart);
    hipLaunchKernelGGL(kern, grid, dim3(tune->block), 0, st, hdr, items, citem, cstart, info->nchunks, info->chunk);
    TT_LAUNCH_CHECK("tt_batch_trace launch");
    // the rays handed over (items with a status array), after the whole z-march
    hipLaunchKernelGGL(trace_time_batch_kernel, dim3((unsigned)info->nchunks), dim3(256), 0, st, items, citem, cstart,
                       info->chunk);
    TT_LAUNCH_CHECK("tt_batch_trace time-pass launch");
    return TT_OK;
}
