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
// rays into a private LDS copy of the bin window its rays touch when that
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
constexpr int kLdsBins = 8192;  // private window: 32 KB of uint32/float

__global__ __launch_bounds__(kDetBlock) void detect_kernel(DetArgs a) {
    __shared__ uint32_t lds[kLdsBins];
    __shared__ int win[4];  // bx0, bx1, by0, by1 of this block's rays
    const int64_t i = (int64_t)blockIdx.x * kDetBlock + threadIdx.x;
    const bool live = i < a.n;
    double x0 = 0, th0 = 0, y0 = 0, ph0 = 0, phase = 0;
    if (live) {
        x0 = a.rf[i];
        th0 = a.rf[a.ld + i];
        y0 = a.rf[2 * a.ld + i];
        ph0 = a.rf[3 * a.ld + i];
        if (a.phase) phase = a.phase[i];
    }
    for (int k = 0; k < a.ndet; ++k) {
        const DetDev &d = a.d[k];
        double x = x0, th = th0, y = y0, ph = ph0;
        int bx = -1, by = -1;
        float w = 1.0f;
        if (live && apply_train(d, x, th, y, ph)) {
            bx = bin_of(x, d.xmin, d.xmax, d.xstep, d.nbx);
            by = bin_of(y, d.ymin, d.ymax, d.ystep, d.nby);
            if (d.mode == 1) w = (float)(1.0 + cos(phase - d.tilt_k * x));
        }
        const bool hit = bx >= 0 && by >= 0;
        // window of bins this block touches
        if (threadIdx.x == 0) {
            win[0] = 0x7fffffff;
            win[1] = -1;
            win[2] = 0x7fffffff;
            win[3] = -1;
        }
        __syncthreads();
        if (hit) {
            atomicMin(&win[0], bx);
            atomicMax(&win[1], bx);
            atomicMin(&win[2], by);
            atomicMax(&win[3], by);
        }
        __syncthreads();
        const int wx0 = win[0], wx1 = win[1], wy0 = win[2], wy1 = win[3];
        const int64_t area = (wx1 >= wx0) ? (int64_t)(wx1 - wx0 + 1) * (wy1 - wy0 + 1) : 0;
        if (area > 0 && area <= kLdsBins) {
            const int ww = wx1 - wx0 + 1;
            for (int j = threadIdx.x; j < area; j += kDetBlock) lds[j] = 0u;
            __syncthreads();
            if (hit) {
                const int j = (by - wy0) * ww + (bx - wx0);
                if (d.mode == 1)
                    atomicAdd(reinterpret_cast<float *>(&lds[j]), w);
                else
                    atomicAdd(&lds[j], 1u);
            }
            __syncthreads();
            for (int j = threadIdx.x; j < area; j += kDetBlock) {
                const uint32_t v = lds[j];
                if (v == 0u) continue;
                const int64_t g = (int64_t)(wy0 + j / ww) * d.nbx + (wx0 + j % ww);
                if (d.mode == 1)
                    atomicAdd(&d.wsum[g], __uint_as_float(v));
                else
                    atomicAdd(&d.counts[g], v);
            }
        } else if (hit) {
            const int64_t g = (int64_t)by * d.nbx + bx;
            if (d.mode == 1)
                atomicAdd(&d.wsum[g], w);
            else
                atomicAdd(&d.counts[g], 1u);
        }
        __syncthreads();
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
    hipLaunchKernelGGL(detect_kernel, dim3((unsigned)tt::ceil_div(n, kDetBlock)), dim3(kDetBlock), 0,
                       tt::as_stream(stream), a);
    TT_LAUNCH_CHECK("tt_detect launch");
    return TT_OK;
}
