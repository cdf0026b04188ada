// Straight-line (Born-series) ray estimate and surrogate features.
//
// Not a reference function: the cheap estimate whose residual against the
// full trace (tt_trace_rk4, the work of particle_tracker.py:312-380) the
// surrogate MLP learns.  Each ray is followed along its unperturbed straight
// line; in every z-cell the node field (tt_grad_stencil layout) is evaluated
// once, at the cell's mid-plane (the field is trilinear, so along a nearly
// axial line the z-integrand is linear inside a cell and the midpoint rule is
// close to exact).  Per ray it accumulates, in z-cell units:
//   S   = int G dZ                 first-order velocity change (x, y, z)
//   M   = int (Zend - Z) G dZ      first-order exit displacement (x, y)
//   T   = int (dG/dx D1x + dG/dy D1y) dZ, R = int (Zend - Z)(...) dZ
//         second-order terms, D1 = first-order displacement along the path
//   P   = int (Phi + dPhi/dx D1x + dPhi/dy D1y) |v|/W dZ   phase, second order
// and the per-segment sums of Gx, Gy, Phi over 16 equal z-segments.
// Outputs: the second-order estimate base = [x, theta, y, phi, phase] (the
// rf layout plus phase, float64) and 48 features per ray (float32).
#include "tt_internal.h"

namespace {

constexpr int kSeg = 16;

struct StraightArgs {
    const float4 *field;
    int nx, ny, nz;
    double x0, y0, z0, dx, dy, dz;
    const double *s0;
    int64_t ld, n;
    float *feat;
    int64_t ldf;
    double *base;  // 5 x ldb
    int64_t ldb;
};

__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float t) {
    return make_float4(fmaf(t, b.x - a.x, a.x), fmaf(t, b.y - a.y, a.y), fmaf(t, b.z - a.z, a.z),
                       fmaf(t, b.w - a.w, a.w));
}

// value and the x/y derivatives (per cell) of channels x, y and w (Phi)
__device__ __forceinline__ void eval_d(const StraightArgs &a, float px, float py, int iz, float tz, float4 &g,
                                       float3 &dgx, float3 &dgy) {
    const bool in = px >= 0.f && px <= (float)(a.nx - 1) && py >= 0.f && py <= (float)(a.ny - 1);
    if (!in) {
        g = make_float4(0.f, 0.f, 0.f, 0.f);
        dgx = dgy = make_float3(0.f, 0.f, 0.f);
        return;
    }
    const int ix = min((int)px, a.nx - 2), iy = min((int)py, a.ny - 2);
    const float tx = px - (float)ix, ty = py - (float)iy;
    const int64_t plane = (int64_t)a.nx * a.ny;
    const float4 *p = a.field + ((int64_t)iz * a.ny + iy) * a.nx + ix;
    const float4 q00 = lerp4(p[0], p[plane], tz), q10 = lerp4(p[1], p[plane + 1], tz);
    const float4 q01 = lerp4(p[a.nx], p[plane + a.nx], tz), q11 = lerp4(p[a.nx + 1], p[plane + a.nx + 1], tz);
    const float4 r0 = lerp4(q00, q10, tx), r1 = lerp4(q01, q11, tx);
    g = lerp4(r0, r1, ty);
    // d/dtx = lerp(q10 - q00, q11 - q01, ty), d/dty = r1 - r0
    dgx = make_float3(fmaf(ty, (q11.x - q01.x) - (q10.x - q00.x), q10.x - q00.x),
                      fmaf(ty, (q11.y - q01.y) - (q10.y - q00.y), q10.y - q00.y),
                      fmaf(ty, (q11.w - q01.w) - (q10.w - q00.w), q10.w - q00.w));
    dgy = make_float3(r1.x - r0.x, r1.y - r0.y, r1.w - r0.w);
}

__global__ __launch_bounds__(256) void straight_kernel(StraightArgs a) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.n) return;
    const double *s = a.s0 + r;
    const double x = s[0], y = s[a.ld], z = s[2 * a.ld];
    const double U0 = s[3 * a.ld] / tt::kC, V0 = s[4 * a.ld] / tt::kC, W0 = s[5 * a.ld] / tt::kC;
    float *f = a.feat + r * a.ldf;
    double *b = a.base + r;
    const double nanv = __longlong_as_double(0x7ff8000000000000ll);
    if (!(W0 > 0.0) || isnan(x) || isnan(y) || isnan(z) || isnan(U0) || isnan(V0)) {
        for (int c = 0; c < 5; ++c) b[c * a.ldb] = nanv;
        for (int c = 0; c < 48; ++c) f[c] = 0.f;
        return;
    }
    const double sx = a.dz / a.dx, sy = a.dz / a.dy;
    const double slx = sx * U0 / W0, sly = sy * V0 / W0;  // cells per z-cell
    double px0 = (x - a.x0) / a.dx, py0 = (y - a.y0) / a.dy, Zs = (z - a.z0) / a.dz;
    if (Zs < 0.0) {  // zero field before the grid
        px0 += slx * -Zs;
        py0 += sly * -Zs;
        Zs = 0.0;
    }
    const double Zend = (double)(a.nz - 1);
    const float iW0 = (float)(1.0 / W0);
    const float sp = (float)(sqrt(U0 * U0 + V0 * V0 + W0 * W0) / W0);
    const float fsx = (float)sx * iW0 * iW0, fsy = (float)sy * iW0 * iW0;  // D1 scale: cells per (int int G)

    float Sx = 0.f, Sy = 0.f, Sz = 0.f, Mx = 0.f, My = 0.f, P = 0.f;
    float Tx = 0.f, Ty = 0.f, Rx = 0.f, Ry = 0.f;
    float Dx = 0.f, Dy = 0.f;  // int int G over the cells so far (first-order displacement / fs)
    float segx = 0.f, segy = 0.f;
    const int ncell = a.nz - 1;
    const int iz0 = Zs >= Zend ? ncell : (int)Zs;
    for (int q = 0; q < 2 * kSeg; ++q) f[q] = 0.f;  // segments the ray never reaches stay 0
    for (int iz = iz0; iz < ncell; ++iz) {
        const double zl = fmax(Zs, (double)iz);
        const float w = (float)((double)(iz + 1) - zl);
        const double zm = zl + 0.5 * (double)w;
        const float px = (float)(px0 + slx * (zm - Zs)), py = (float)(py0 + sly * (zm - Zs));
        float4 g;
        float3 dgx, dgy;
        eval_d(a, px, py, iz, (float)(zm - iz), g, dgx, dgy);
        const float h = 0.5f * w;
        // first-order displacement (per fs) at the midpoint
        const float d1x = fsx * (Dx + Sx * h + 0.5f * g.x * h * h);
        const float d1y = fsy * (Dy + Sy * h + 0.5f * g.y * h * h);
        const float f2x = dgx.x * d1x + dgy.x * d1y;  // dGx/dx D1x + dGx/dy D1y
        const float f2y = dgx.y * d1x + dgy.y * d1y;
        const float lever = (float)(Zend - zm);
        Tx = fmaf(w, f2x, Tx);
        Ty = fmaf(w, f2y, Ty);
        Rx = fmaf(w * lever, f2x, Rx);
        Ry = fmaf(w * lever, f2y, Ry);
        Mx = fmaf(w * lever, g.x, Mx);
        My = fmaf(w * lever, g.y, My);
        Dx += Sx * w + 0.5f * g.x * w * w;
        Dy += Sy * w + 0.5f * g.y * w * w;
        Sx = fmaf(w, g.x, Sx);
        Sy = fmaf(w, g.y, Sy);
        Sz = fmaf(w, g.z, Sz);
        P = fmaf(w * sp, g.w + dgx.z * d1x + dgy.z * d1y, P);  // phase to second order
        segx = fmaf(w, g.x, segx);
        segy = fmaf(w, g.y, segy);
        // segment q holds cells [q ncell / 16, (q + 1) ncell / 16)
        const int seg = (iz * kSeg) / ncell;
        if (((iz + 1) * kSeg) / ncell != seg || iz + 1 == ncell) {
            f[seg] = segx;
            f[kSeg + seg] = segy;
            segx = segy = 0.f;
        }
    }
    // second-order estimate at the exit plane
    const double Ue = U0 + (double)(Sx + Tx) * iW0, Ve = V0 + (double)(Sy + Ty) * iW0;
    const double We = W0 + (double)Sz * iW0;
    const double L = Zend - Zs;
    const double pxe = px0 + slx * L + (double)fsx * (double)(Mx + Rx);
    const double pye = py0 + sly * L + (double)fsy * (double)(My + Ry);
    b[0] = a.x0 + pxe * a.dx;
    b[a.ldb] = atan(Ue / We);
    b[2 * a.ldb] = a.y0 + pye * a.dy;
    b[3 * a.ldb] = atan(Ve / We);
    b[4 * a.ldb] = (double)P;
    // global features
    float *gf = f + 2 * kSeg;
    gf[0] = Sx;
    gf[1] = Sy;
    gf[2] = Mx;
    gf[3] = My;
    gf[4] = Tx;
    gf[5] = Ty;
    gf[6] = Rx;
    gf[7] = Ry;
    gf[8] = Sz;
    gf[9] = P;
    gf[10] = (float)U0;
    gf[11] = (float)V0;
    gf[12] = (float)(W0 - 1.0);
    gf[13] = (float)(px0 - floor(px0));
    gf[14] = (float)(py0 - floor(py0));
    gf[15] = (float)(Zend - Zs);
}

// corrected = base + Y[:, :5] * scale; flag when a predicted error bound
// (Y[:, 5] * scale_err[0] for position, Y[:, 6] * scale_err[1] for angle)
// exceeds its threshold, or when the base estimate is not finite.  Flagged ray
// indices are appended to list (count in *nflag).
__global__ void apply_kernel(const double *base, int64_t ldb, const float *Y, int64_t ldy, int64_t n, float s_pos,
                             float s_ang, float s_ph, float e_pos, float e_ang, double t_pos, double t_ang,
                             double *rf, int64_t ldr, double *phase, int32_t *flag, int32_t *list, int32_t *nflag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
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
    if (bad && list) list[atomicAdd(nflag, 1)] = (int32_t)i;
}

}  // namespace

extern "C" int tt_surrogate_apply(const double *base, int64_t ldb, const float *Y, int64_t ldy, int64_t n,
                                  const float *scales, const double *thresholds, double *rf, int64_t ldr,
                                  double *phase, int32_t *flag, int32_t *list, int32_t *nflag, void *stream) {
    TT_REQUIRE(base && Y && scales && thresholds && rf, "tt_surrogate_apply: null pointer");
    TT_REQUIRE(n >= 0 && ldb >= n && ldr >= n && ldy >= 7, "tt_surrogate_apply: bad sizes");
    TT_REQUIRE(!list || nflag, "tt_surrogate_apply: list needs nflag");
    if (n == 0) return TT_OK;
    hipLaunchKernelGGL(apply_kernel, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0, tt::as_stream(stream), base,
                       ldb, Y, ldy, n, scales[0], scales[1], scales[2], scales[3], scales[4], thresholds[0],
                       thresholds[1], rf, ldr, phase, flag, list, nflag);
    TT_LAUNCH_CHECK("tt_surrogate_apply launch");
    return TT_OK;
}

extern "C" int tt_straight_features(const float *field4, const tt_grid *grid, const double *s0, int64_t ld, int64_t n,
                                    float *feat, int64_t ldf, double *base, int64_t ldb, void *stream) {
    TT_REQUIRE(field4 && grid && s0 && feat && base, "tt_straight_features: null pointer");
    TT_REQUIRE(n >= 0 && ld >= n && ldb >= n && ldf >= 48, "tt_straight_features: bad sizes");
    TT_REQUIRE(grid->nx >= 2 && grid->ny >= 2 && grid->nz >= 2, "tt_straight_features: grid needs >= 2 nodes per axis");
    TT_REQUIRE(grid->dx > 0 && grid->dy > 0 && grid->dz > 0, "tt_straight_features: spacing must be > 0");
    if (n == 0) return TT_OK;
    StraightArgs a;
    a.field = reinterpret_cast<const float4 *>(field4);
    a.nx = grid->nx;
    a.ny = grid->ny;
    a.nz = grid->nz;
    a.x0 = grid->x0;
    a.y0 = grid->y0;
    a.z0 = grid->z0;
    a.dx = grid->dx;
    a.dy = grid->dy;
    a.dz = grid->dz;
    a.s0 = s0;
    a.ld = ld;
    a.n = n;
    a.feat = feat;
    a.ldf = ldf;
    a.base = base;
    a.ldb = ldb;
    hipLaunchKernelGGL(straight_kernel, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0, tt::as_stream(stream), a);
    TT_LAUNCH_CHECK("tt_straight_features launch");
    return TT_OK;
}
