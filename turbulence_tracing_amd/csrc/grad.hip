// Gradient stencil: the work of ElectronCube.calc_dndr
// (particle_tracking/particle_tracker.py:220-241) as one HBM-bound pass.
//
// For each node: ne/nc = alpha + beta f, clipped at ne_max (PT:230-231);
// -0.5 d(ne/nc)/dx_i by np.gradient's rule (central interior, one-sided
// first-order edges; PT:235-237); the result is written as one float4 per
// node {Gx, Gy, Gz, Phi} in the x-fastest/z-slowest layout the tracer reads:
//   G   = -0.5 dz grad(ne/nc)          (velocity change per z-cell)
//   Phi = k0 dz (sqrt(1 - ne/nc) - 1)  (phase per z-cell)
// Algorithmic traffic: sizeof(T) read + 16 B written per node (neighbour
// re-reads hit L1/L2).
#include <algorithm>

#include "tt_internal.h"

namespace {

template <typename T>
struct GradArgs {
    const T *f;
    int64_t sx, sy, sz;
    int nx, ny, nz;
    T beta, clipv;  // s = min(beta f, clipv); ne/nc = alpha + s
    double alpha;
    T cx, cy, cz;   // -0.5 dz / (2 d_i): scale of a central difference
    float phase_scale;  // k0 dz
    const double *sumsq; // optional: beta /= sqrt(*sumsq / count) (unit-RMS normalisation)
    double count;
    float4 *out;
};

template <typename T>
__device__ __forceinline__ T load_s(const GradArgs<T> &a, T beta, int ix, int iy, int iz) {
    const T v = a.f[(int64_t)ix * a.sx + (int64_t)iy * a.sy + (int64_t)iz * a.sz] * beta;
    return v < a.clipv ? v : a.clipv;
}

// one-dimensional np.gradient at index i of n given the stencil values
template <typename T>
__device__ __forceinline__ T grad1(T sm, T s0, T sp, int i, int n, T c) {
    if (i == 0) return (sp - s0) * (c + c);
    if (i == n - 1) return (s0 - sm) * (c + c);
    return (sp - sm) * c;
}

// A thread walks one (x, y) column through kZRun consecutive z-planes,
// carrying the z-neighbours in registers: per node one new load along the
// column and four x / y neighbour loads that hit in L1 (the block's 64 x 4
// columns read the same plane together), and one float4 store.
constexpr int kZRun = 16;

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T>
__global__ __launch_bounds__(256) void grad_kernel(GradArgs<T> a) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    const int z0 = blockIdx.z * kZRun;
    if (ix >= a.nx || iy >= a.ny) return;
    const int z1 = min(z0 + kZRun, a.nz);
    T beta = a.beta;
    if (a.sumsq) beta = (T)((double)beta / sqrt(*a.sumsq / a.count));
    T zm = z0 > 0 ? load_s(a, beta, ix, iy, z0 - 1) : (T)0;
    T s0 = load_s(a, beta, ix, iy, z0);
    for (int iz = z0; iz < z1; ++iz) {
        __builtin_amdgcn_s_setprio(1);  // the plane's loads leave before other waves' arithmetic
        const T zp = iz < a.nz - 1 ? load_s(a, beta, ix, iy, iz + 1) : s0;
        const T xm = ix > 0 ? load_s(a, beta, ix - 1, iy, iz) : s0;
        const T xp = ix < a.nx - 1 ? load_s(a, beta, ix + 1, iy, iz) : s0;
        const T ym = iy > 0 ? load_s(a, beta, ix, iy - 1, iz) : s0;
        const T yp = iy < a.ny - 1 ? load_s(a, beta, ix, iy + 1, iz) : s0;
        __builtin_amdgcn_s_setprio(0);
        float4 o;
        o.x = (float)grad1(xm, s0, xp, ix, a.nx, a.cx);
        o.y = (float)grad1(ym, s0, yp, iy, a.ny, a.cy);
        o.z = (float)grad1(zm, s0, zp, iz, a.nz, a.cz);
        // n - 1 = -r / (1 + sqrt(1 - r)), evaluated without cancellation
        const double r = a.alpha + (double)s0;
        const double q = 1.0 - r;
        const double eta = q > 0.0 ? -r / (1.0 + sqrt(q)) : -1.0;
        o.w = (float)(eta * (double)a.phase_scale);
        // streamed to HBM (nontemporal): the field is far larger than the caches
        const f32x4 ov = {o.x, o.y, o.z, o.w};
        __builtin_nontemporal_store(ov, reinterpret_cast<f32x4 *>(a.out + ((int64_t)iz * a.ny + iy) * a.nx + ix));
        zm = s0;
        s0 = zp;
    }
}

// Plane forms of the node field for the tracer's cell cache: node (ix, iy,
// iz) gets the bilinear form of the z-plane cell whose lower corner it is,
//   A = f00, B = f10 - f00, C = f01 - f00, E = (f11 - f01) - B,
// so that f(tx, ty) = A + B tx + C ty + E tx ty on that plane (the last
// column and row, which start no cell, get A = f and B = C = E = 0).  The
// subtractions are those the tracer would do per cell, so the forms equal
// the ones it forms itself, bit for bit.  One thread per output float4
// (component j = t % 4 of node t / 4), so a wave's stores are one contiguous
// 1 KiB run; the neighbour loads hit L1 / L2.  Algorithmic traffic: 16 B read
// + 64 B written per node.
__device__ __forceinline__ void plane_form(const float4 *__restrict__ f, float4 *__restrict__ out, int nx, int ny,
                                           int64_t t) {
    const int j = (int)(t & 3);
    const int64_t i = t >> 2;
    const int ix = (int)(i % nx);
    const int iy = (int)((i / nx) % ny);
    const float4 a = f[i];
    float4 o = a;
    if (j > 0) {
        o = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ix < nx - 1 && iy < ny - 1) {
            const float4 f10 = f[i + 1], f01 = f[i + nx];
            const float4 b = make_float4(f10.x - a.x, f10.y - a.y, f10.z - a.z, f10.w - a.w);
            if (j == 1) {
                o = b;
            } else if (j == 2) {
                o = make_float4(f01.x - a.x, f01.y - a.y, f01.z - a.z, f01.w - a.w);
            } else {
                const float4 f11 = f[i + nx + 1];
                o = make_float4((f11.x - f01.x) - b.x, (f11.y - f01.y) - b.y, (f11.z - f01.z) - b.z,
                                (f11.w - f01.w) - b.w);
            }
        }
    }
    // streamed to HBM (nontemporal): written once, read by the tracer later
    __builtin_nontemporal_store((f32x4){o.x, o.y, o.z, o.w}, reinterpret_cast<f32x4 *>(out + t));
}

__global__ __launch_bounds__(256) void plane_forms_kernel(const float4 *__restrict__ f, float4 *__restrict__ out,
                                                          int nx, int ny, int64_t n4) {
    // grid-stride: a launch holds fewer than 2^32 work-items
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n4; t += (int64_t)gridDim.x * blockDim.x)
        plane_form(f, out, nx, ny, t);
}

template <typename T>
int launch(const void *field, int64_t sx, int64_t sy, int64_t sz, const tt_grid *g, double alpha, double beta,
           double ne_max, double k0, const double *sumsq, int64_t count, float *out4, hipStream_t st) {
    GradArgs<T> a;
    a.f = static_cast<const T *>(field);
    a.sx = sx;
    a.sy = sy;
    a.sz = sz;
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.beta = (T)beta;
    a.clipv = (T)(ne_max - alpha);
    a.alpha = alpha;
    a.cx = (T)(-0.5 * g->dz / (2.0 * g->dx));
    a.cy = (T)(-0.5 * g->dz / (2.0 * g->dy));
    a.cz = (T)(-0.5 * g->dz / (2.0 * g->dz));
    a.phase_scale = (float)(k0 * g->dz);
    a.sumsq = sumsq;
    a.count = (double)count;
    a.out = reinterpret_cast<float4 *>(out4);
    dim3 block(64, 4, 1);
    dim3 grid((unsigned)tt::ceil_div(g->nx, 64), (unsigned)tt::ceil_div(g->ny, 4), (unsigned)tt::ceil_div(g->nz, kZRun));
    hipLaunchKernelGGL(grad_kernel<T>, grid, block, 0, st, a);
    TT_LAUNCH_CHECK("tt_grad_stencil launch");
    return TT_OK;
}

}  // namespace

extern "C" int tt_grad_stencil(const void *field, int32_t dtype, int64_t sx, int64_t sy, int64_t sz,
                               const tt_grid *grid, double alpha, double beta, double ne_max, double k0,
                               const double *sumsq, int64_t count, float *out4, void *stream) {
    TT_REQUIRE(field && grid && out4, "tt_grad_stencil: null pointer");
    TT_REQUIRE(grid->nx >= 2 && grid->ny >= 2 && grid->nz >= 2,
               "tt_grad_stencil: every axis needs >= 2 nodes (np.gradient), got %d x %d x %d", grid->nx,
               grid->ny, grid->nz);
    TT_REQUIRE(grid->nz <= 65535, "tt_grad_stencil: nz > 65535");
    TT_REQUIRE(grid->dx > 0 && grid->dy > 0 && grid->dz > 0, "tt_grad_stencil: spacing must be > 0");
    TT_REQUIRE(!sumsq || count > 0, "tt_grad_stencil: count must be > 0 with sumsq");
    hipStream_t st = tt::as_stream(stream);
    if (dtype == 0) return launch<float>(field, sx, sy, sz, grid, alpha, beta, ne_max, k0, sumsq, count, out4, st);
    if (dtype == 1) return launch<double>(field, sx, sy, sz, grid, alpha, beta, ne_max, k0, sumsq, count, out4, st);
    tt::set_error("tt_grad_stencil: dtype must be 0 (float32) or 1 (float64), got %d", dtype);
    return TT_ERR_ARG;
}

extern "C" int tt_plane_forms(const float *field4, const tt_grid *grid, float *forms, void *stream) {
    TT_REQUIRE(field4 && grid && forms, "tt_plane_forms: null pointer");
    TT_REQUIRE(grid->nx >= 2 && grid->ny >= 2 && grid->nz >= 1, "tt_plane_forms: grid needs >= 2 nodes in x and y");
    TT_REQUIRE(forms != field4, "tt_plane_forms: forms must not alias field4");
    const int64_t n4 = 4 * (int64_t)grid->nx * grid->ny * grid->nz;
    const int64_t blocks = std::min<int64_t>(tt::ceil_div(n4, 256), (int64_t)1 << 22);
    hipLaunchKernelGGL(plane_forms_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       tt::as_stream(stream), reinterpret_cast<const float4 *>(field4),
                       reinterpret_cast<float4 *>(forms), grid->nx, grid->ny, n4);
    TT_LAUNCH_CHECK("tt_plane_forms launch");
    return TT_OK;
}
