// Detectors: the ray-transfer optics and Rays.histogram of
// particle_tracking/ray_transfer_matrix.py (RTM:37-154, 173-191, 208-299).
//
// Each ray's [x, theta, y, phi] goes through the optical train in fp64
// registers (a 4x4 block-diagonal matrix per element is two scalar updates;
// rejected rays become NaN, RTM:77-78), then into one or more histograms.
// Binning follows np.histogram2d exactly: edges e_i = i*step + min with
// step = (max - min)/n (np.linspace), bin = searchsorted(e, v, 'right') - 1,
// v == max counts in the last bin, outside values and NaN are dropped.
// Histograms are stored transposed, H[iy][ix] (RTM:191).
//
// Counts are uint32 atomics (order-independent, bit-reproducible);
// interferogram weights are float32 atomics.  Each workgroup first bins its
// 1024 rays into a private LDS copy of the bin window they touch when that
// window is small (rays sorted by entry column land close together on the
// detector), then flushes the non-zero bins; otherwise it adds to global
// memory directly.
#include "tt_internal.h"

#pragma clang fp contract(off)

namespace {

struct DetDev {
    int32_t n_ops, to_mm, mode, pad;
    tt_optic ops[TT_MAX_OPS];
    int32_t nbx, nby;
    double xmin, xmax, ymin, ymax, xstep, ystep;
    double tilt_k;
    uint32_t *counts;
    float *wsum;
};

struct DetArgs {
    DetDev d[TT_MAX_DETECTORS];
    int32_t ndet;
    const double *rf;
    int64_t ld, n;
    const double *phase;
};

// apply the train; returns false when the ray is rejected (or NaN)
__device__ __forceinline__ bool apply_train(const DetDev &d, double &x, double &th, double &y, double &ph) {
    if (d.to_mm) {
        x = x * 1e3;
        y = y * 1e3;
    }
    for (int i = 0; i < d.n_ops; ++i) {
        const tt_optic o = d.ops[i];
        switch (o.op) {
            case TT_OP_DISTANCE:
                x = x + o.a * th;
                y = y + o.a * ph;
                break;
            case TT_OP_LENS:
                th = (-1.0 / o.a) * x + th;
                ph = (-1.0 / o.b) * y + ph;
                break;
            case TT_OP_APERTURE:
                if (x * x + y * y > o.a * o.a) return false;
                break;
            case TT_OP_STOP:
                if (x * x + y * y < o.a * o.a) return false;
                break;
            case TT_OP_KNIFE: {
                const double v = o.axis == 2 ? y : x;
                if ((o.b > 0 && v > o.a) || (o.b < 0 && v < o.a)) return false;
                break;
            }
            case TT_OP_RECT:
                if (x * x > o.a * o.a && y * y > o.b * o.b) return false;
                break;
            case TT_OP_ANNULUS: {
                const double r2 = x * x + y * y;
                if (r2 > o.a * o.a && r2 < o.b * o.b) return false;
                break;
            }
            default:
                break;
        }
    }
    return !(isnan(x) || isnan(th) || isnan(y) || isnan(ph));
}

// np.histogram(bins=n, range=(lo, hi)) bin of v, or -1 if dropped
__device__ __forceinline__ int bin_of(double v, double lo, double hi, double step, int n) {
    if (!(v >= lo && v <= hi)) return -1;
    if (v == hi) return n - 1;
    int b = (int)floor((v - lo) / step);
    b = b < 0 ? 0 : (b > n - 1 ? n - 1 : b);
    // correct against the exact np.linspace edges (i*step + lo; last = hi)
    while (b > 0 && v < (double)b * step + lo) --b;
    while (b < n - 1 && v >= ((b + 1 == n) ? hi : (double)(b + 1) * step + lo)) ++b;
    return b;
}

constexpr int kDetBlock = 256;
constexpr int kDetRays = 4;                       // rays per thread
constexpr int kDetBlockRays = kDetBlock * kDetRays;
constexpr int kLdsBins = 8192;  // private window: 32 KB of uint32/float

// LDS-only workgroup barrier: the fences order LDS operations only, so the
// barrier does not wait for this wave's global atomics to complete
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// A block bins 1024 rays (4 per thread, coalesced) into each detector in
// turn: the bin window of its hits (wave reductions, then 4 LDS atomics per
// wave), a private LDS copy of that window when it holds at most kLdsBins
// bins, then the non-zero bins flushed with global atomics.  The barriers
// order LDS only, so the flush of one detector overlaps the next.
__global__ __launch_bounds__(kDetBlock) void detect_kernel(DetArgs a) {
    __shared__ uint32_t lds[kLdsBins];
    __shared__ int win[4];  // bx0, bx1, by0, by1 of this block's hits
    double x0[kDetRays], th0[kDetRays], y0[kDetRays], ph0[kDetRays], phase[kDetRays];
    bool live[kDetRays];
#pragma unroll
    for (int j = 0; j < kDetRays; ++j) {
        const int64_t i = (int64_t)blockIdx.x * kDetBlockRays + j * kDetBlock + threadIdx.x;
        live[j] = i < a.n;
        x0[j] = th0[j] = y0[j] = ph0[j] = phase[j] = 0.0;
        if (live[j]) {
            x0[j] = a.rf[i];
            th0[j] = a.rf[a.ld + i];
            y0[j] = a.rf[2 * a.ld + i];
            ph0[j] = a.rf[3 * a.ld + i];
            if (a.phase) phase[j] = a.phase[i];
        }
    }
    if (threadIdx.x == 0) {
        win[0] = 0x7fffffff;
        win[1] = -1;
        win[2] = 0x7fffffff;
        win[3] = -1;
    }
    for (int k = 0; k < a.ndet; ++k) {
        const DetDev &d = a.d[k];
        int bx[kDetRays], by[kDetRays];
        float w[kDetRays];
        int lx0 = 0x7fffffff, lx1 = -1, ly0 = 0x7fffffff, ly1 = -1;
#pragma unroll
        for (int j = 0; j < kDetRays; ++j) {
            double x = x0[j], th = th0[j], y = y0[j], ph = ph0[j];
            bx[j] = by[j] = -1;
            w[j] = 1.0f;
            if (live[j] && apply_train(d, x, th, y, ph)) {
                bx[j] = bin_of(x, d.xmin, d.xmax, d.xstep, d.nbx);
                by[j] = bin_of(y, d.ymin, d.ymax, d.ystep, d.nby);
                if (d.mode == 1) w[j] = (float)(1.0 + cos(phase[j] - d.tilt_k * x));
            }
            if (bx[j] < 0 || by[j] < 0) bx[j] = by[j] = -1;
            if (bx[j] >= 0) {
                lx0 = min(lx0, bx[j]);
                lx1 = max(lx1, bx[j]);
                ly0 = min(ly0, by[j]);
                ly1 = max(ly1, by[j]);
            }
        }
        lx0 = wave_min(lx0);
        lx1 = wave_max(lx1);
        ly0 = wave_min(ly0);
        ly1 = wave_max(ly1);
        lds_barrier();  // win is initialised; the previous detector's LDS flush reads are done
        if ((threadIdx.x & 63) == 0 && lx1 >= 0) {
            atomicMin(&win[0], lx0);
            atomicMax(&win[1], lx1);
            atomicMin(&win[2], ly0);
            atomicMax(&win[3], ly1);
        }
        lds_barrier();
        const int wx0 = win[0], wx1 = win[1], wy0 = win[2], wy1 = win[3];
        const int64_t area = (wx1 >= wx0) ? (int64_t)(wx1 - wx0 + 1) * (wy1 - wy0 + 1) : 0;
        if (area > 0 && area <= kLdsBins) {
            const int ww = wx1 - wx0 + 1;
            for (int j = threadIdx.x; j < area; j += kDetBlock) lds[j] = 0u;
            lds_barrier();
#pragma unroll
            for (int j = 0; j < kDetRays; ++j) {
                if (bx[j] < 0) continue;
                const int q = (by[j] - wy0) * ww + (bx[j] - wx0);
                if (d.mode == 1)
                    atomicAdd(reinterpret_cast<float *>(&lds[q]), w[j]);
                else
                    atomicAdd(&lds[q], 1u);
            }
            lds_barrier();
            for (int j = threadIdx.x; j < area; j += kDetBlock) {
                const uint32_t v = lds[j];
                if (v == 0u) continue;
                const int64_t gi = (int64_t)(wy0 + j / ww) * d.nbx + (wx0 + j % ww);
                if (d.mode == 1)
                    atomicAdd(&d.wsum[gi], __uint_as_float(v));
                else
                    atomicAdd(&d.counts[gi], v);
            }
        } else {
#pragma unroll
            for (int j = 0; j < kDetRays; ++j) {
                if (bx[j] < 0) continue;
                const int64_t gi = (int64_t)by[j] * d.nbx + bx[j];
                if (d.mode == 1)
                    atomicAdd(&d.wsum[gi], w[j]);
                else
                    atomicAdd(&d.counts[gi], 1u);
            }
        }
        lds_barrier();  // every thread has read win
        if (threadIdx.x == 0) {
            win[0] = 0x7fffffff;
            win[1] = -1;
            win[2] = 0x7fffffff;
            win[3] = -1;
        }
    }
}

__global__ __launch_bounds__(256) void optics_kernel(DetDev d, const double *rin, int64_t ldi, int64_t n,
                                                     double *rout, int64_t ldo) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = rin[i], th = rin[ldi + i], y = rin[2 * ldi + i], ph = rin[3 * ldi + i];
    const bool ok = apply_train(d, x, th, y, ph);
    const double nanv = __longlong_as_double(0x7ff8000000000000ll);
    rout[i] = ok ? x : nanv;
    rout[ldo + i] = ok ? th : nanv;
    rout[2 * ldo + i] = ok ? y : nanv;
    rout[3 * ldo + i] = ok ? ph : nanv;
}

int fill_train(const tt_train *t, DetDev &d) {
    TT_REQUIRE(t, "null optical train");
    TT_REQUIRE(t->n_ops >= 0 && t->n_ops <= TT_MAX_OPS, "optical train has %d ops (max %d)", t->n_ops, TT_MAX_OPS);
    d.n_ops = t->n_ops;
    d.to_mm = t->to_mm;
    for (int i = 0; i < t->n_ops; ++i) {
        const tt_optic &o = t->ops[i];
        TT_REQUIRE(o.op >= TT_OP_DISTANCE && o.op <= TT_OP_ANNULUS, "optical op %d: unknown code %d", i, o.op);
        if (o.op == TT_OP_LENS) TT_REQUIRE(o.a != 0 && o.b != 0, "optical op %d: lens focal length 0", i);
        if (o.op == TT_OP_KNIFE) {
            TT_REQUIRE(o.axis == 0 || o.axis == 2, "optical op %d: knife edge axis must be 0 (x) or 2 (y)", i);
            TT_REQUIRE(o.b != 0, "optical op %d: knife edge direction must be <0 or >0", i);
        }
        d.ops[i] = o;
    }
    return TT_OK;
}

}  // namespace

extern "C" int tt_optics_apply(const tt_train *train, const double *r_in, int64_t ld_in, int64_t n, double *r_out,
                               int64_t ld_out, void *stream) {
    DetDev d{};
    int rc = fill_train(train, d);
    if (rc) return rc;
    TT_REQUIRE(n >= 0 && ld_in >= n && ld_out >= n, "tt_optics_apply: bad sizes");
    if (n == 0) return TT_OK;
    TT_REQUIRE(r_in && r_out, "tt_optics_apply: null pointer");
    hipLaunchKernelGGL(optics_kernel, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0, tt::as_stream(stream), d,
                       r_in, ld_in, n, r_out, ld_out);
    TT_LAUNCH_CHECK("tt_optics_apply launch");
    return TT_OK;
}

extern "C" int tt_detect(const tt_detector *dets, int32_t ndet, const double *rf, int64_t ld, const double *phase,
                         int64_t n, void *stream) {
    TT_REQUIRE(dets && ndet >= 1 && ndet <= TT_MAX_DETECTORS, "tt_detect: need 1..%d detectors", TT_MAX_DETECTORS);
    TT_REQUIRE(n >= 0 && ld >= n, "tt_detect: bad sizes");
    if (n == 0) return TT_OK;
    TT_REQUIRE(rf, "tt_detect: null rays");
    DetArgs a{};
    for (int k = 0; k < ndet; ++k) {
        const tt_detector &s = dets[k];
        DetDev &d = a.d[k];
        int rc = fill_train(&s.train, d);
        if (rc) return rc;
        const tt_hist &h = s.hist;
        TT_REQUIRE(h.nbx >= 1 && h.nby >= 1, "tt_detect: detector %d has no bins", k);
        TT_REQUIRE(h.xmax > h.xmin && h.ymax > h.ymin, "tt_detect: detector %d has an empty range", k);
        TT_REQUIRE(s.mode == 0 || s.mode == 1, "tt_detect: detector %d mode must be 0 or 1", k);
        TT_REQUIRE(s.mode == 0 ? s.counts != nullptr : s.wsum != nullptr, "tt_detect: detector %d output is null", k);
        TT_REQUIRE(s.mode == 0 || phase, "tt_detect: interferogram needs the phase");
        d.mode = s.mode;
        d.nbx = h.nbx;
        d.nby = h.nby;
        d.xmin = h.xmin;
        d.xmax = h.xmax;
        d.ymin = h.ymin;
        d.ymax = h.ymax;
        d.xstep = (h.xmax - h.xmin) / h.nbx;
        d.ystep = (h.ymax - h.ymin) / h.nby;
        d.tilt_k = s.tilt_k;
        d.counts = s.counts;
        d.wsum = s.wsum;
    }
    a.ndet = ndet;
    a.rf = rf;
    a.ld = ld;
    a.n = n;
    a.phase = phase;
    hipLaunchKernelGGL(detect_kernel, dim3((unsigned)tt::ceil_div(n, (int64_t)kDetBlockRays)), dim3(kDetBlock), 0,
                       tt::as_stream(stream), a);
    TT_LAUNCH_CHECK("tt_detect launch");
    return TT_OK;
}
