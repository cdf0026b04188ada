// Beam sampler (the distribution of ElectronCube.init_beam,
// particle_tracking/particle_tracker.py:258-310) and the ray ordering used by
// the tracer (a counting sort by the Morton code of the entry column).
#include <mutex>

#include "tt_internal.h"

namespace {

// PT:273-278 draws t = 2 pi U, u = U + U (reflected at 1), phi = pi U,
// chi = sigma N; here from Philox counters (ray index + offset, stream 1,
// words 0 and 1), converted exactly as oracle/philox.py:beam_uniforms.
__global__ __launch_bounds__(256) void beam_kernel(int64_t n, double beam_size, double divergence, double extent,
                                                   int axis, uint32_t k0, uint32_t k1, uint64_t offset, double *s0,
                                                   int64_t ld) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t idx = (uint64_t)i + offset;
    const uint32_t lo = (uint32_t)idx, hi = (uint32_t)(idx >> 32);
    const tt::u32x4 ra = tt::philox4x32_10(tt::u32x4{lo, hi, tt::kStreamBeam, 0u}, k0, k1);
    const tt::u32x4 rb = tt::philox4x32_10(tt::u32x4{lo, hi, tt::kStreamBeam, 1u}, k0, k1);
    const double t = 2.0 * M_PI * tt::u01d(ra.x);
    double u = tt::u01d(ra.y) + tt::u01d(ra.z);
    if (u > 1.0) u = 2.0 - u;
    const double phi = M_PI * tt::u01d(ra.w);
    const double nrm = sqrt(-2.0 * log(tt::u01d(rb.x))) * cos(2.0 * M_PI * tt::u01d(rb.y));
    const double chi = divergence * nrm;
    const double sc = sin(chi), cc = cos(chi), sp = sin(phi), cp = cos(phi);
    const double pa = beam_size * u * cos(t), pb = beam_size * u * sin(t);
    // PT:280-309: axis 0 (x) starts at +extent, axes 1 and 2 at -extent; the
    // along-axis velocity is c cos(chi), the two transverse ones
    // c sin(chi) cos(phi) and c sin(chi) sin(phi).  Written as per-row selects
    // (a branchy local array of six doubles was mis-lowered for gfx950).
    const double vl = tt::kC * cc, va = tt::kC * sc * cp, vb = tt::kC * sc * sp;
    const bool ax0 = axis == 0, ax1 = axis == 1;
    double *o = s0 + i;
    o[0] = ax0 ? extent : pa;
    o[ld] = ax0 ? pa : (ax1 ? -extent : pb);
    o[2 * ld] = ax0 ? pb : (ax1 ? pb : -extent);
    o[3 * ld] = ax0 ? vl : va;
    o[4 * ld] = ax0 ? va : (ax1 ? vl : vb);
    o[5 * ld] = ax0 ? vb : (ax1 ? vb : vl);
}

// ---------------------------------------------------------------- ray sort
struct SortGeom {
    double x0, y0, idx, idy;
    int nx, ny, shift, bits;
};

__device__ __forceinline__ uint32_t spread16(uint32_t v) {
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ __forceinline__ uint32_t ray_key(const SortGeom &g, double x, double y) {
    double fx = (x - g.x0) * g.idx, fy = (y - g.y0) * g.idy;
    int ix = isnan(fx) ? 0 : (int)fmin(fmax(floor(fx), 0.0), (double)(g.nx - 1));
    int iy = isnan(fy) ? 0 : (int)fmin(fmax(floor(fy), 0.0), (double)(g.ny - 1));
    return spread16((uint32_t)(ix >> g.shift)) | (spread16((uint32_t)(iy >> g.shift)) << 1);
}

__global__ void fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void count_kernel(const double *s0, int64_t ld, int64_t n, SortGeom g,
                                                    uint32_t *hist) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    atomicAdd(&hist[ray_key(g, s0[i], s0[ld + i])], 1u);
}

constexpr int kScanBlock = 1024;
constexpr int kScanItems = 4;
constexpr int kScanChunk = kScanBlock * kScanItems;

// exclusive scan of one chunk in place; chunk total to sums[blockIdx.x]
__global__ __launch_bounds__(kScanBlock) void scan_chunk(uint32_t *data, int64_t n, uint32_t *sums) {
    __shared__ uint32_t warp_tot[kScanBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = (base + j < n) ? data[base + j] : 0u;
        const uint32_t t = v[j];
        v[j] = run;
        run += t;
    }
    // inclusive scan of the per-thread totals across the wavefront
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = run;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    if (lane == 63) warp_tot[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < kScanBlock / 64; ++w) {
            const uint32_t t = warp_tot[w];
            warp_tot[w] = acc;
            acc += t;
        }
        if (sums) sums[blockIdx.x] = acc;
    }
    __syncthreads();
    const uint32_t excl = incl - run + warp_tot[wave];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < n) data[base + j] = v[j] + excl;
}

__global__ __launch_bounds__(kScanBlock) void add_chunk(uint32_t *data, int64_t n, const uint32_t *sums) {
    const uint32_t add = sums[blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    for (int j = threadIdx.x; j < kScanChunk; j += kScanBlock)
        if (base + j < n) data[base + j] += add;
}

__global__ __launch_bounds__(256) void scatter_kernel(const double *s0, int64_t ld, int64_t n, SortGeom g,
                                                      uint32_t *offsets, int32_t *order) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t pos = atomicAdd(&offsets[ray_key(g, s0[i], s0[ld + i])], 1u);
    order[pos] = (int32_t)i;
}

// The scatter that also moves the rays: each ray's six values go, as one
// 64-byte record, to its sorted position in a ray-major staging buffer
// (three 16-byte stores into one aligned record instead of six scattered
// 8-byte ones), and unpack_kernel then writes the sorted rows coalesced.
__global__ __launch_bounds__(256) void scatter_rays_kernel(const double *s0, int64_t ld, int64_t n, SortGeom g,
                                                           uint32_t *offsets, int32_t *order, double *staging) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = s0[i], y = s0[ld + i], z = s0[2 * ld + i];
    const double vx = s0[3 * ld + i], vy = s0[4 * ld + i], vz = s0[5 * ld + i];
    const uint32_t pos = atomicAdd(&offsets[ray_key(g, x, y)], 1u);
    order[pos] = (int32_t)i;
    double2 *rec = reinterpret_cast<double2 *>(staging + (int64_t)pos * 8);
    rec[0] = make_double2(x, y);
    rec[1] = make_double2(z, vx);
    rec[2] = make_double2(vy, vz);
}

__global__ __launch_bounds__(256) void unpack_kernel(const double *staging, int64_t n, double *out, int64_t ldo) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 *rec = reinterpret_cast<const double2 *>(staging + i * 8);
    const double2 a = rec[0], b = rec[1], c = rec[2];
    out[i] = a.x;
    out[ldo + i] = a.y;
    out[2 * ldo + i] = b.x;
    out[3 * ldo + i] = b.y;
    out[4 * ldo + i] = c.x;
    out[5 * ldo + i] = c.y;
}

int sort_geom(const tt_grid *grid, int shift, SortGeom &g, int64_t &nbins) {
    TT_REQUIRE(grid && grid->nx >= 1 && grid->ny >= 1, "tt_ray_sort: bad grid");
    TT_REQUIRE(shift >= 0 && shift <= 15, "tt_ray_sort: brick_shift must be in [0, 15]");
    TT_REQUIRE(grid->dx > 0 && grid->dy > 0, "tt_ray_sort: spacing must be > 0");
    g.x0 = grid->x0;
    g.y0 = grid->y0;
    g.idx = 1.0 / grid->dx;
    g.idy = 1.0 / grid->dy;
    g.nx = grid->nx;
    g.ny = grid->ny;
    g.shift = shift;
    const int cols = ((grid->nx > grid->ny ? grid->nx : grid->ny) - 1) >> shift;
    int bits = 0;
    while ((1 << bits) <= cols) ++bits;
    TT_REQUIRE(bits <= 12, "tt_ray_sort: too many columns; raise brick_shift");
    g.bits = bits;
    nbins = (int64_t)1 << (2 * bits);
    return TT_OK;
}

}  // namespace

extern "C" int tt_beam_sample(int64_t n, double beam_size, double divergence, double extent, int32_t axis,
                              uint64_t seed, uint64_t offset, double *s0, int64_t ld, void *stream) {
    TT_REQUIRE(n >= 0 && (n == 0 || s0), "tt_beam_sample: bad arguments");
    TT_REQUIRE(ld >= n, "tt_beam_sample: ld < n");
    TT_REQUIRE(axis >= 0 && axis <= 2, "tt_beam_sample: axis must be 0, 1 or 2");
    if (n == 0) return TT_OK;
    hipLaunchKernelGGL(beam_kernel, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0, tt::as_stream(stream), n,
                       beam_size, divergence, extent, axis, (uint32_t)seed, (uint32_t)(seed >> 32), offset, s0, ld);
    TT_LAUNCH_CHECK("tt_beam_sample launch");
    return TT_OK;
}

extern "C" int tt_ray_sort_scratch_bytes(const tt_grid *grid, int32_t brick_shift, size_t *bytes) {
    SortGeom g;
    int64_t nbins;
    int rc = sort_geom(grid, brick_shift, g, nbins);
    if (rc) return rc;
    TT_REQUIRE(bytes, "tt_ray_sort_scratch_bytes: null");
    const int64_t chunks = tt::ceil_div(nbins, kScanChunk);
    *bytes = (size_t)(nbins + 2 * chunks + 64) * sizeof(uint32_t);
    return TT_OK;
}

namespace {
int sort_impl(const double *s0, int64_t ld, int64_t n, const tt_grid *grid, int32_t brick_shift, void *scratch,
              int32_t *order, double *staging, double *out, int64_t ldo, void *stream) {
    SortGeom g;
    int64_t nbins;
    int rc = sort_geom(grid, brick_shift, g, nbins);
    if (rc) return rc;
    TT_REQUIRE(n >= 0 && n < ((int64_t)1 << 31), "tt_ray_sort: n must be < 2^31");
    TT_REQUIRE(ld >= n, "tt_ray_sort: ld < n");
    if (n == 0) return TT_OK;
    TT_REQUIRE(s0 && scratch && order, "tt_ray_sort: null pointer");
    hipStream_t st = tt::as_stream(stream);
    uint32_t *hist = static_cast<uint32_t *>(scratch);
    const int64_t chunks = tt::ceil_div(nbins, kScanChunk);
    uint32_t *sums = hist + nbins;
    uint32_t *sums2 = sums + chunks;
    TT_REQUIRE(chunks <= kScanChunk, "tt_ray_sort: too many bins");
    hipLaunchKernelGGL(fill_u32, dim3((unsigned)tt::ceil_div(nbins, 256)), dim3(256), 0, st, hist, nbins, 0u);
    hipLaunchKernelGGL(count_kernel, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0, st, s0, ld, n, g, hist);
    hipLaunchKernelGGL(scan_chunk, dim3((unsigned)chunks), dim3(kScanBlock), 0, st, hist, nbins, sums);
    hipLaunchKernelGGL(scan_chunk, dim3(1), dim3(kScanBlock), 0, st, sums, chunks, sums2);
    hipLaunchKernelGGL(add_chunk, dim3((unsigned)chunks), dim3(kScanBlock), 0, st, hist, nbins,
                       (const uint32_t *)sums);
    const dim3 nb((unsigned)tt::ceil_div(n, 256));
    if (out) {
        hipLaunchKernelGGL(scatter_rays_kernel, nb, dim3(256), 0, st, s0, ld, n, g, hist, order, staging);
        hipLaunchKernelGGL(unpack_kernel, nb, dim3(256), 0, st, (const double *)staging, n, out, ldo);
    } else {
        hipLaunchKernelGGL(scatter_kernel, nb, dim3(256), 0, st, s0, ld, n, g, hist, order);
    }
    TT_LAUNCH_CHECK("tt_ray_sort launch");
    return TT_OK;
}
}  // namespace

extern "C" int tt_ray_sort(const double *s0, int64_t ld, int64_t n, const tt_grid *grid, int32_t brick_shift,
                           void *scratch, int32_t *order, void *stream) {
    return sort_impl(s0, ld, n, grid, brick_shift, scratch, order, nullptr, nullptr, 0, stream);
}

// ------------------------------------------------- allocation-free ordering
// tt_ray_sort_noalloc keeps its column counts and scan partials in this
// file-scope array instead of a caller-provided scratch buffer: room for
// 2^20 entry columns (Morton keys of up to 1024 x 1024 columns, i.e. grids up
// to 1024 nodes per axis at brick_shift 0).  Each device holds its own copy
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device.
constexpr int64_t kStaticSortBins = (int64_t)1 << 20;
constexpr int64_t kStaticSortWords = kStaticSortBins + 2 * (kStaticSortBins / kScanChunk) + 64;
__device__ __align__(256) uint32_t scratch[kStaticSortWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// the address of `scratch` on the current device, looked up once per device
int static_sort_scratch(void **out) {
    int dev = -1;
    TT_HIP(hipGetDevice(&dev), "tt_ray_sort_noalloc: hipGetDevice");
    TT_REQUIRE(dev >= 0 && dev < kMaxDevices, "tt_ray_sort_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        TT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "tt_ray_sort_noalloc: hipGetSymbolAddress");
        TT_REQUIRE(ptr, "tt_ray_sort_noalloc: null symbol address");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return TT_OK;
}
}  // namespace

extern "C" int tt_ray_sort_noalloc(const double *s0, int64_t ld, int64_t n, const tt_grid *grid,
                                   int32_t brick_shift, int32_t *order, void *stream) {
    SortGeom g;
    int64_t nbins;
    int rc = sort_geom(grid, brick_shift, g, nbins);
    if (rc) return rc;
    TT_REQUIRE(nbins <= kStaticSortBins,
               "tt_ray_sort_noalloc: %lld entry-column bins exceed the static scratch of %lld (grids up to "
               "1024 columns per axis after brick_shift); use tt_ray_sort with a scratch buffer",
               (long long)nbins, (long long)kStaticSortBins);
    TT_REQUIRE(n >= 0 && ld >= n, "tt_ray_sort_noalloc: bad sizes");
    if (n == 0) return TT_OK;  // an empty ray set: nothing is read or written
    void *ptr = nullptr;
    rc = static_sort_scratch(&ptr);
    if (rc) return rc;
    // sort_impl clears the counts with a kernel before counting, and the scan
    // writes every partial before it is read
    return sort_impl(s0, ld, n, grid, brick_shift, ptr, order, nullptr, nullptr, 0, stream);
}

extern "C" int tt_ray_sort_rays(const double *s0, int64_t ld, int64_t n, const tt_grid *grid, int32_t brick_shift,
                                void *scratch, int32_t *order, double *staging, double *out, int64_t ldo,
                                void *stream) {
    TT_REQUIRE(ldo >= n, "tt_ray_sort_rays: ldo < n");
    TT_REQUIRE(n == 0 || (staging && out && ((uintptr_t)staging & 15) == 0),
               "tt_ray_sort_rays: staging must be 16-byte aligned");
    TT_REQUIRE(n == 0 || out != s0, "tt_ray_sort_rays: out must not alias s0");
    return sort_impl(s0, ld, n, grid, brick_shift, scratch, order, staging, out, ldo, stream);
}
