// Gaussian random field by inverse 3-D FFT: the work of gaussian3D_FFT
// (gaussian_fields/turboGen.py:488-538), restructured for the GPU.
//
// The spectrum F(k) = [Wr(c) + Wr(c') + i (Wi(c) - Wi(c'))] sqrt(P(k)) is
// Hermitian (c' is the centred-grid partner of c, the np.flip of TG:525-526
// for odd sizes and the index of -k for any size), so the real output
// ifftn(F).real is produced from the half spectrum kx in [0, nx/2]:
//   pass Z: generate F (noise x sqrt P, zero mode, TG:522-534) for a block of
//           kx columns directly into LDS and inverse-FFT along z;
//   pass Y: inverse-FFT along y;
//   pass X: complete each row by Hermitian symmetry, inverse-FFT along x,
//           keep the real part times 1/(nx ny nz) (numpy's ifftn scale) and
//           reduce the sum of squares for the unit-RMS normalisation.
// Each pass stages a batch of pencils in LDS and runs a Stockham autosort
// FFT (radix 8/4/2/3/5 butterflies, direct DFT passes for other primes)
// with the twiddle table in LDS.  The spectrum F is never written: pass Z
// writes the z-transformed half spectrum.  Traffic per field:
// 8 (Z) + 16 (Y) + 8 + 4 (X) bytes per half-spectrum element.
#include <climits>
#include <cmath>
#include <vector>

#include "tt_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kElems = 32;                     // complex values per thread per pass
constexpr int kBatch = kThreads * kElems;      // complex values per block
constexpr int kMaxRad = 24;
constexpr int kMaxBins = 2048;  // power-spectrum bins (two doubles each in LDS)

struct AxisPlan {
    int n, nrad;
    int rad[kMaxRad];
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 mul_i(float2 a) { return make_float2(-a.y, a.x); }  // * (+i)

// inverse-sign (e^{+2 pi i q r / R}) DFT of R points in registers
template <int R>
__device__ __forceinline__ void dft(float2 *v);

template <>
__device__ __forceinline__ void dft<2>(float2 *v) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft<4>(float2 *v) {
    const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const float2 s13 = cadd(v[1], v[3]), d13 = mul_i(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}

template <>
__device__ __forceinline__ void dft<8>(float2 *v) {
    // split into even/odd radix-4 DFTs, then combine with e^{+2 pi i q / 8}
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft<4>(e);
    dft<4>(o);
    const float h = 0.70710678118654752f;
    const float2 w1 = cmul(o[1], make_float2(h, h));
    const float2 w2 = mul_i(o[2]);
    const float2 w3 = cmul(o[3], make_float2(-h, h));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
}

template <>
__device__ __forceinline__ void dft<3>(float2 *v) {
    const float c = -0.5f, s = 0.86602540378443865f;  // e^{2 pi i / 3}
    const float2 a = v[0], b = v[1], d = v[2];
    const float2 t = cadd(b, d), u = csub(b, d);
    const float2 m = make_float2(a.x + c * t.x, a.y + c * t.y);
    const float2 su = make_float2(-s * u.y, s * u.x);  // i s (b - d)
    v[0] = cadd(a, t);
    v[1] = cadd(m, su);
    v[2] = csub(m, su);
}

template <>
__device__ __forceinline__ void dft<5>(float2 *v) {
    const float c1 = 0.30901699437494742f, s1 = 0.95105651629515357f;   // 2 pi / 5
    const float c2 = -0.80901699437494742f, s2 = 0.58778525229247313f;  // 4 pi / 5
    const float2 a = v[0];
    const float2 t1 = cadd(v[1], v[4]), u1 = csub(v[1], v[4]);
    const float2 t2 = cadd(v[2], v[3]), u2 = csub(v[2], v[3]);
    const float2 m1 = make_float2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    const float2 m2 = make_float2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    // i (s1 u1 + s2 u2) and i (s2 u1 - s1 u2)
    const float2 n1 = make_float2(-(s1 * u1.y + s2 * u2.y), s1 * u1.x + s2 * u2.x);
    const float2 n2 = make_float2(-(s2 * u1.y - s1 * u2.y), s2 * u1.x - s1 * u2.x);
    v[0] = make_float2(a.x + t1.x + t2.x, a.y + t1.y + t2.y);
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
}

// One Stockham pass of radix R over B pencils of length n in LDS (pencil p
// at buf + p*stride).  Every thread first computes all of its butterflies
// into registers, then all write: in-place with two barriers.
template <int R>
__device__ __forceinline__ void pass_r_body(float2 *buf, int n, int stride, int B, int Ns, const float2 *tw) {
    constexpr int NB = (kElems + R - 1) / R;
    float2 v[NB][R];
    const int nbf = n / R;
    const int total = B * nbf;
    const int twstep = n / (Ns * R);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int g = threadIdx.x + i * kThreads;
        if (g < total) {
            const int p = g / nbf, b = g - p * nbf;
            const int k = b % Ns;
            const float2 *pb = buf + p * stride;
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = pb[b + r * nbf];
            if (k) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], tw[r * k * twstep]);
            }
            dft<R>(v[i]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int g = threadIdx.x + i * kThreads;
        if (g < total) {
            const int p = g / nbf, b = g - p * nbf;
            const int k = b % Ns;
            float2 *pb = buf + p * stride + (b / Ns) * Ns * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) pb[q * Ns] = v[i][q];
        }
    }
    __syncthreads();
}

template <int R>
__device__ __noinline__ void pass_r(float2 *buf, int n, int stride, int B, int Ns, const float2 *tw) {
    pass_r_body<R>(buf, n, stride, B, Ns, tw);
}

// Direct DFT pass for any radix R (odd primes > 5): one output per item.
__device__ __noinline__ void pass_generic(float2 *buf, int n, int stride, int B, int Ns, int R, const float2 *tw) {
    float2 o[kElems];
    const int nbf = n / R;
    const int total = B * n;
    const int twstep = n / (Ns * R);
#pragma unroll
    for (int i = 0; i < kElems; ++i) {
        const int g = threadIdx.x + i * kThreads;
        o[i] = make_float2(0.f, 0.f);
        if (g < total) {
            const int p = g / n, e = g - p * n;
            const int q = e / nbf, b = e - q * nbf;
            const int k = b % Ns;
            const float2 *pb = buf + p * stride;
            float2 acc = make_float2(0.f, 0.f);
            int qr = 0;  // (q * r) mod R
            for (int r = 0; r < R; ++r) {
                float2 x = pb[b + r * nbf];
                if (k) x = cmul(x, tw[r * k * twstep]);
                const float2 w = tw[qr * nbf];
                acc.x += x.x * w.x - x.y * w.y;
                acc.y += x.x * w.y + x.y * w.x;
                qr += q;
                if (qr >= R) qr -= R;
            }
            o[i] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kElems; ++i) {
        const int g = threadIdx.x + i * kThreads;
        if (g < total) {
            const int p = g / n, e = g - p * n;
            const int q = e / nbf, b = e - q * nbf;
            const int k = b % Ns;
            buf[p * stride + (b / Ns) * Ns * R + k + q * Ns] = o[i];
        }
    }
    __syncthreads();
}

__device__ void lds_fft(float2 *buf, int stride, int B, const AxisPlan &ap, const float2 *tw) {
    int Ns = 1;
    for (int s = 0; s < ap.nrad; ++s) {
        const int R = ap.rad[s];
        switch (R) {
            case 8: pass_r<8>(buf, ap.n, stride, B, Ns, tw); break;
            case 4: pass_r<4>(buf, ap.n, stride, B, Ns, tw); break;
            case 2: pass_r<2>(buf, ap.n, stride, B, Ns, tw); break;
            case 3: pass_r<3>(buf, ap.n, stride, B, Ns, tw); break;
            case 5: pass_r<5>(buf, ap.n, stride, B, Ns, tw); break;
            default: pass_generic(buf, ap.n, stride, B, Ns, R, tw); break;
        }
        Ns *= R;
    }
}

// ---------------------------------------------------------------------
// Fixed lengths (128, 192, 256, 384, 512, 1024: powers of two times at most
// one 3) with every size known at compile time, so the index arithmetic is shifts and masks, and with the first and
// the last Stockham pass in registers: the first pass reads its butterfly
// inputs straight from the pass's loader (global memory or the generator)
// and the last one hands its outputs straight to the storer, so a pencil
// goes through LDS twice fewer than with LDS-to-LDS passes only.  Same radix
// order as factorize (8 while 8 divides what is left, then 4 or 2, then the
// odd prime).
template <int N, int NS>
constexpr int radix_at() {
    int m = N / NS, e = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++e;
    }
    if (e >= 3) return 8;
    if (e == 2) return 4;
    if (e == 1) return 2;
    for (int p = 3;; p += 2)
        if (m % p == 0) return p;
}

// butterfly beta of B * NB -> (pencil p, index b); PFAST: p varies fastest
// (pencils are columns of a row-major plane), else b (pencils are rows)
template <int B, int NB, bool PFAST>
__device__ __forceinline__ void split_bf(int beta, int &p, int &b) {
    if (PFAST) {
        p = beta % B;
        b = beta / B;
    } else {
        p = beta / NB;
        b = beta % NB;
    }
}

// first pass (Ns = 1): inputs are elements b + NB r of pencil p
template <int N, bool PFAST, class Load>
__device__ __forceinline__ void first_pass(float2 *buf, Load load) {
    constexpr int R = radix_at<N, 1>(), NB = N / R, B = kBatch / N, TOT = B * NB;
    constexpr int PER = (TOT + kThreads - 1) / kThreads;
    float2 v[PER][R];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int beta = threadIdx.x + i * kThreads;
        if (TOT % kThreads == 0 || beta < TOT) {
            int p, b;
            split_bf<B, NB, PFAST>(beta, p, b);
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = load(p, b + NB * r);
            dft<R>(v[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int beta = threadIdx.x + i * kThreads;
        if (TOT % kThreads == 0 || beta < TOT) {
            int p, b;
            split_bf<B, NB, PFAST>(beta, p, b);
#pragma unroll
            for (int q = 0; q < R; ++q) buf[p * (N + 1) + R * b + q] = v[i][q];
        }
    }
    __syncthreads();
}

// last pass (Ns = N / R): outputs are elements b + NS q of pencil p
template <int N, int NS, bool PFAST, class Store>
__device__ __forceinline__ void last_pass(const float2 *buf, const float2 *tw, Store store) {
    constexpr int R = N / NS, B = kBatch / N, TOT = B * NS;
    constexpr int PER = (TOT + kThreads - 1) / kThreads;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int beta = threadIdx.x + i * kThreads;
        if (TOT % kThreads == 0 || beta < TOT) {
            int p, b;
            split_bf<B, NS, PFAST>(beta, p, b);
            float2 v[R];
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = buf[p * (N + 1) + b + NS * r];
            if (b) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[r * b]);
            }
            dft<R>(v);
#pragma unroll
            for (int q = 0; q < R; ++q) store(p, b + NS * q, v[q]);
        }
    }
}

template <int N, int NS, bool PFAST, class Store>
__device__ __forceinline__ void rest_passes(float2 *buf, const float2 *tw, Store store) {
    constexpr int R = radix_at<N, NS>();
    if constexpr (NS * R == N) {
        last_pass<N, NS, PFAST>(buf, tw, store);
    } else {
        pass_r_body<R>(buf, N, N + 1, kBatch / N, NS, tw);
        rest_passes<N, NS * R, PFAST>(buf, tw, store);
    }
}

// the whole transform of the block's B pencils: load(p, j) -> store(p, j, v)
template <int N, bool PFAST, class Load, class Store>
__device__ __forceinline__ void fft_pow2(float2 *buf, const float2 *tw, Load load, Store store) {
    first_pass<N, PFAST>(buf, load);
    rest_passes<N, radix_at<N, 1>(), PFAST>(buf, tw, store);
}

enum { kGenNative = 0, kGenNoise = 1, kPassY = 2, kPassX = 3, kFwdX = 4, kPowZ = 5 };

struct PassArgs {
    AxisPlan ap;
    const float2 *tw;  // twiddles of this axis, e^{+2 pi i m / n}
    int nx, ny, nz, nxh;
    int B;
    float2 *spec;  // [nz][ny][nxh]
    // native spectrum
    double aniso, kmin, kmax;
    float half_beta, sqrt_amp;
    uint32_t key0, key1;
    // caller noise
    const float *wr, *wi, *sp;
    // output
    float *out;
    float scale;
    double *sumsq;
    // power spectrum (kFwdX input, kPowZ binning)
    const float *in;
    int mode, nbins;
    double fz_step, fy_step, fx_step;  // 1 / (n dx) per axis (spectrum_3D_scalar frequencies)
    double bin_w;                      // bin width in |k| (spectrum_3D_scalar)
    double pscale;                     // factor on |F|^2 (compute3Dspectrum: 1 / nt^2)
    double *hsum, *hcnt;
};

// (j + n/2) mod n and (2 (n/2) - c) mod n for j, c in [0, n): one conditional
// subtraction each instead of an integer division
__device__ __forceinline__ int centred(int j, int n) {
    const int v = j + n / 2;
    return v >= n ? v - n : v;
}
__device__ __forceinline__ int partner(int c, int n) {
    const int v = 2 * (n / 2) - c;
    return v >= n ? v - n : v;
}

// np.fft.fftfreq(n)[j] = j_signed * (1.0 / n)
__device__ __forceinline__ double fftfreq(int j, int n) {
    const int js = (j <= (n - 1) / 2) ? j : j - n;
    return (double)js * (1.0 / (double)n);
}

__device__ __forceinline__ float2 philox_pair(int64_t lin, uint32_t k0, uint32_t k1) {
    const tt::u32x4 r =
        tt::philox4x32_10(tt::u32x4{(uint32_t)lin, (uint32_t)((uint64_t)lin >> 32), tt::kStreamField, 0u}, k0, k1);
    const float u1 = tt::u01f(r.x), u2 = tt::u01f(r.y);
    // Box-Muller on the hardware log2, sqrt and sin / cos (v_sin / v_cos take
    // revolutions, so sin(2 pi u2) is v_sin(u2)); u1, u2 lie in (0, 1) and are
    // never denormal.  Within ~1e-6 of the correctly rounded forms, which cost
    // 15% of the field's time
    const float rad = __builtin_amdgcn_sqrtf(-2.0f * 0.69314718055994531f * __builtin_amdgcn_logf(u1));
    return make_float2(rad * __builtin_amdgcn_cosf(u2), rad * __builtin_amdgcn_sinf(u2));
}

#pragma clang fp contract(off)
__device__ float sqrt_psd_native(const PassArgs &a, int jx, int jy, int jz) {
    const double fx = fftfreq(jx, a.nx), fy = fftfreq(jy, a.ny), fz = fftfreq(jz, a.nz) / a.aniso;
    const double Ka = sqrt((fx * fx + fy * fy) + fz * fz);
    if (!(Ka >= a.kmin && Ka <= a.kmax && Ka > 0.0)) return 0.0f;
    return a.sqrt_amp * __builtin_amdgcn_exp2f(a.half_beta * __builtin_amdgcn_logf((float)Ka));  // Ka^(beta/2)
}
#pragma clang fp contract(on)

template <int MODE>
__device__ __forceinline__ float2 gen(const PassArgs &a, int jx, int jy, int jz) {
    if ((jx | jy | jz) == 0) return make_float2(0.f, 0.f);  // TG:534
    const int cx = centred(jx, a.nx), cy = centred(jy, a.ny), cz = centred(jz, a.nz);
    const int px = partner(cx, a.nx), py = partner(cy, a.ny), pz = partner(cz, a.nz);
    const int64_t lc = ((int64_t)cx * a.ny + cy) * a.nz + cz;
    const int64_t lp = ((int64_t)px * a.ny + py) * a.nz + pz;
    float2 wc, wp;
    float s;
    if (MODE == kGenNative) {
        s = sqrt_psd_native(a, jx, jy, jz);
        if (s == 0.0f) return make_float2(0.f, 0.f);
        wc = philox_pair(lc, a.key0, a.key1);
        wp = philox_pair(lp, a.key0, a.key1);
    } else {
        s = a.sp[lc];
        wc = make_float2(a.wr[lc], a.wi[lc]);
        wp = make_float2(a.wr[lp], a.wi[lp]);
    }
    return make_float2((wc.x + wp.x) * s, (wc.y - wp.y) * s);
}

// ---------------------------------------------------------------------
// Power spectrum of a real field (the estimators of
// gaussian_fields/calculate_spectrum_3d.py:3-59 and cmpspec.py:147-194).
// The passes above run with the inverse sign; for a real field that gives
// conj(fftn(f)), whose |.|^2 is fftn's.  The field is read as rows along x'
// (the reference's last axis, contiguous in C order), so x' / y' / z' are
// the reference's axes 2 / 1 / 0.  Each stored half-spectrum element also
// stands for its mirror (-k), which has the same power, unless kx' is 0 or
// nx'/2.
// Mode 0 (spectrum_3D_scalar): K = sqrt((k0^2 + k1^2) + k2^2) with k_i the
//   fftfreq(M_i, dx) value of axis i; bin b holds bin_w (b) <= K <
//   bin_w (b + 1), the reference's comparisons, for b <= nbins - 2.
// Mode 1 (compute3Dspectrum): shell round(sqrt(v0^2 + v1^2 + v2^2)) of the
//   reference's loop values v_i in range(-M_i // 2, M_i // 2 - 1), so one
//   index per axis, (M_i - 2) // 2, is never visited; power scaled by
//   1 / nt^2 (rh = fftn(r) / nt).
__device__ __forceinline__ int signed_of(int j, int n) { return (j <= (n - 1) / 2) ? j : j - n; }
// the reference loop value of array index j, or INT_MIN when not visited
__device__ __forceinline__ int loop_value(int j, int n) {
    const int hi = n / 2 - 2;
    if (j == (n - 2) / 2) return INT_MIN;
    return j <= hi ? j : j - n;
}

#pragma clang fp contract(off)
__device__ __forceinline__ void bin_power(const PassArgs &a, int jx, int jy, int jz, float power, double *hs,
                                          double *hc) {
    const bool self_conj = jx == 0 || (a.nx % 2 == 0 && jx == a.nx / 2);
    if (a.mode == 0) {
        const double f0 = (double)signed_of(jz, a.nz) * a.fz_step;
        const double f1 = (double)signed_of(jy, a.ny) * a.fy_step;
        const double f2 = (double)signed_of(jx, a.nx) * a.fx_step;
        const double K = sqrt((f0 * f0 + f1 * f1) + f2 * f2);
        int b = (int)(K / a.bin_w);
        if (K < (double)b * a.bin_w) --b;
        else if (K >= (double)(b + 1) * a.bin_w) ++b;
        if (b < 0 || b > a.nbins - 2) return;
        const double w = self_conj ? 1.0 : 2.0;
        atomicAdd(&hs[b], w * (double)power);
        atomicAdd(&hc[b], w);
        return;
    }
    const double p = (double)power * a.pscale;
    for (int m = 0; m < (self_conj ? 1 : 2); ++m) {
        const int ix = m ? (a.nx - jx) % a.nx : jx, iy = m ? (a.ny - jy) % a.ny : jy, iz = m ? (a.nz - jz) % a.nz : jz;
        const int v0 = loop_value(iz, a.nz), v1 = loop_value(iy, a.ny), v2 = loop_value(ix, a.nx);
        if (v0 == INT_MIN || v1 == INT_MIN || v2 == INT_MIN) continue;
        const double rk = sqrt((double)v0 * v0 + (double)v1 * v1 + (double)v2 * v2);
        const int sh = (int)rint(rk);
        if (sh < a.nbins) {
            atomicAdd(&hs[sh], p);
            atomicAdd(&hc[sh], 1.0);
        }
    }
}
#pragma clang fp contract(on)

// NF > 0: the pencil length is NF (one of the fixed lengths) and B = kBatch / NF, both
// compile-time constants; NF == 0: any length, from the plan.
template <int MODE, int NF>
__global__ __launch_bounds__(kThreads) void fft_pass_kernel(PassArgs a) {
    extern __shared__ float2 smem[];
    const int n = NF ? NF : a.ap.n;
    const int stride = n + 1;
    const int B = NF ? kBatch / NF : a.B;
    float2 *buf = smem;
    float2 *tw = smem + B * stride;
    for (int j = threadIdx.x; j < n; j += kThreads) tw[j] = a.tw[j];
    const int total = B * n;

    if constexpr (NF > 0) {
        constexpr int BF = kBatch / NF;
        if constexpr (MODE == kPowZ) {
            double *hs = reinterpret_cast<double *>(tw + NF), *hc = hs + a.nbins;
            for (int j = threadIdx.x; j < 2 * a.nbins; j += kThreads) hs[j] = 0.0;
            const int kx0 = blockIdx.x * BF, fixed = blockIdx.y;
            fft_pow2<NF, true>(
                buf, tw,
                [&](int p, int j) {
                    const int kx = kx0 + p;
                    return kx < a.nxh ? a.spec[((int64_t)j * a.ny + fixed) * a.nxh + kx] : make_float2(0.f, 0.f);
                },
                [&](int p, int j, float2 v) {
                    const int kx = kx0 + p;
                    if (kx < a.nxh) bin_power(a, kx, fixed, j, v.x * v.x + v.y * v.y, hs, hc);
                });
            __syncthreads();
            for (int j = threadIdx.x; j < a.nbins; j += kThreads) {
                if (hc[j] != 0.0) {
                    atomicAdd(&a.hsum[j], hs[j]);
                    atomicAdd(&a.hcnt[j], hc[j]);
                }
            }
        } else if constexpr (MODE == kFwdX) {
            const int64_t rows = (int64_t)a.ny * a.nz, r0 = (int64_t)blockIdx.x * BF;
            fft_pow2<NF, false>(
                buf, tw,
                [&](int p, int k) { return make_float2(r0 + p < rows ? a.in[(r0 + p) * NF + k] : 0.f, 0.f); },
                [&](int p, int k, float2 v) {
                    if (r0 + p < rows && k < a.nxh) a.spec[(r0 + p) * a.nxh + k] = v;
                });
        } else if constexpr (MODE == kPassX) {
            const int64_t rows = (int64_t)a.ny * a.nz, r0 = (int64_t)blockIdx.x * BF;
            double ss = 0.0;
            fft_pow2<NF, false>(
                buf, tw,
                [&](int p, int k) {
                    if (r0 + p >= rows) return make_float2(0.f, 0.f);
                    const float2 *srow = a.spec + (r0 + p) * a.nxh;
                    if (k < a.nxh) return srow[k];
                    const float2 c = srow[NF - k];
                    return make_float2(c.x, -c.y);
                },
                [&](int p, int k, float2 v) {
                    if (r0 + p >= rows) return;
                    const float o = v.x * a.scale;
                    a.out[(r0 + p) * NF + k] = o;
                    ss += (double)o * (double)o;
                });
            if (a.sumsq) {
                __shared__ double red[kThreads / 64];
                for (int off = 32; off > 0; off >>= 1) ss += __shfl_down(ss, off, 64);
                if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
                __syncthreads();
                if (threadIdx.x == 0) {
                    double t = 0.0;
                    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
                    atomicAdd(a.sumsq, t);
                }
            }
        } else {  // kGenNative, kGenNoise (pencils along z) and kPassY (along y)
            const int kx0 = blockIdx.x * BF, fixed = blockIdx.y;
            fft_pow2<NF, true>(
                buf, tw,
                [&](int p, int j) {
                    const int kx = kx0 + p;
                    if (kx >= a.nxh) return make_float2(0.f, 0.f);
                    if constexpr (MODE == kPassY)
                        return a.spec[((int64_t)fixed * a.ny + j) * a.nxh + kx];
                    else
                        return gen<MODE>(a, kx, fixed, j);
                },
                [&](int p, int j, float2 v) {
                    const int kx = kx0 + p;
                    if (kx >= a.nxh) return;
                    const int64_t idx = (MODE == kPassY) ? ((int64_t)fixed * a.ny + j) * a.nxh + kx
                                                         : ((int64_t)j * a.ny + fixed) * a.nxh + kx;
                    a.spec[idx] = v;
                });
        }
    } else if (MODE == kPowZ) {
        // pencils along z of the half spectrum, then |F|^2 into a block-private
        // histogram (after the FFT's LDS) flushed with global atomics
        double *hs = reinterpret_cast<double *>(tw + n), *hc = hs + a.nbins;
        for (int j = threadIdx.x; j < 2 * a.nbins; j += kThreads) hs[j] = 0.0;
        const int kx0 = blockIdx.x * B;
        const int fixed = blockIdx.y;
        for (int g = threadIdx.x; g < total; g += kThreads) {
            const int p = g % B, j = g / B;
            const int kx = kx0 + p;
            buf[p * stride + j] = kx < a.nxh ? a.spec[((int64_t)j * a.ny + fixed) * a.nxh + kx] : make_float2(0.f, 0.f);
        }
        __syncthreads();
        lds_fft(buf, stride, B, a.ap, tw);
        for (int g = threadIdx.x; g < total; g += kThreads) {
            const int p = g % B, j = g / B;
            const int kx = kx0 + p;
            if (kx >= a.nxh) continue;
            const float2 v = buf[p * stride + j];
            bin_power(a, kx, fixed, j, v.x * v.x + v.y * v.y, hs, hc);
        }
        __syncthreads();
        for (int j = threadIdx.x; j < a.nbins; j += kThreads) {
            if (hc[j] != 0.0) {
                atomicAdd(&a.hsum[j], hs[j]);
                atomicAdd(&a.hcnt[j], hc[j]);
            }
        }
    } else if (MODE == kFwdX) {
        // real rows along x' -> first nx'/2 + 1 outputs of their transform
        const int64_t rows = (int64_t)a.ny * a.nz;
        const int64_t r0 = (int64_t)blockIdx.x * B;
        for (int g = threadIdx.x; g < total; g += kThreads) {
            const int p = g / n, k = g - p * n;
            const int64_t row = r0 + p;
            buf[p * stride + k] = make_float2(row < rows ? a.in[row * n + k] : 0.f, 0.f);
        }
        __syncthreads();
        lds_fft(buf, stride, B, a.ap, tw);
        for (int g = threadIdx.x; g < total; g += kThreads) {
            const int p = g / n, k = g - p * n;
            const int64_t row = r0 + p;
            if (row < rows && k < a.nxh) a.spec[row * a.nxh + k] = buf[p * stride + k];
        }
    } else if (MODE == kGenNative || MODE == kGenNoise || MODE == kPassY) {
        // pencils along z (gen) or y: B consecutive kx columns at fixed ky (gen) / z (Y)
        const int kx0 = blockIdx.x * B;
        const int fixed = blockIdx.y;
        for (int g = threadIdx.x; g < total; g += kThreads) {
            const int p = g % B, j = g / B;
            const int kx = kx0 + p;
            float2 v = make_float2(0.f, 0.f);
            if (kx < a.nxh) {
                if (MODE == kPassY)
                    v = a.spec[((int64_t)fixed * a.ny + j) * a.nxh + kx];
                else
                    v = gen<MODE>(a, kx, fixed, j);
            }
            buf[p * stride + j] = v;
        }
        __syncthreads();
        lds_fft(buf, stride, B, a.ap, tw);
        for (int g = threadIdx.x; g < total; g += kThreads) {
            const int p = g % B, j = g / B;
            const int kx = kx0 + p;
            if (kx >= a.nxh) continue;
            const int64_t idx = (MODE == kPassY) ? ((int64_t)fixed * a.ny + j) * a.nxh + kx
                                                 : ((int64_t)j * a.ny + fixed) * a.nxh + kx;
            a.spec[idx] = buf[p * stride + j];
        }
    } else {
        // rows along x: complete by Hermitian symmetry, transform, keep Re
        const int64_t rows = (int64_t)a.ny * a.nz;
        const int64_t r0 = (int64_t)blockIdx.x * B;
        for (int g = threadIdx.x; g < total; g += kThreads) {
            const int p = g / n, k = g - p * n;
            const int64_t row = r0 + p;
            float2 v = make_float2(0.f, 0.f);
            if (row < rows) {
                const float2 *srow = a.spec + row * a.nxh;
                if (k < a.nxh) {
                    v = srow[k];
                } else {
                    const float2 c = srow[n - k];
                    v = make_float2(c.x, -c.y);
                }
            }
            buf[p * stride + k] = v;
        }
        __syncthreads();
        lds_fft(buf, stride, B, a.ap, tw);
        double ss = 0.0;
        for (int g = threadIdx.x; g < total; g += kThreads) {
            const int p = g / n, k = g - p * n;
            const int64_t row = r0 + p;
            if (row >= rows) continue;
            const float v = buf[p * stride + k].x * a.scale;
            a.out[row * n + k] = v;
            ss += (double)v * (double)v;
        }
        if (a.sumsq) {
            // block reduction of the sum of squares
            __shared__ double red[kThreads / 64];
            for (int off = 32; off > 0; off >>= 1) ss += __shfl_down(ss, off, 64);
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
            __syncthreads();
            if (threadIdx.x == 0) {
                double t = 0.0;
                for (int w = 0; w < kThreads / 64; ++w) t += red[w];
                atomicAdd(a.sumsq, t);
            }
        }
    }
}

void factorize(int n, AxisPlan &ap) {
    ap.n = n;
    ap.nrad = 0;
    int m = n;
    int e2 = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++e2;
    }
    while (e2 >= 3) {
        ap.rad[ap.nrad++] = 8;
        e2 -= 3;
    }
    if (e2 == 2) ap.rad[ap.nrad++] = 4;
    if (e2 == 1) ap.rad[ap.nrad++] = 2;
    for (int p = 3; m > 1; p += 2) {
        while (m % p == 0) {
            ap.rad[ap.nrad++] = p;
            m /= p;
        }
    }
}

int pencils_per_block(int n) {
    int B = kBatch / n;
    return B < 1 ? 1 : B;
}

size_t smem_bytes(int n, int B) { return ((size_t)B * (n + 1) + n) * sizeof(float2); }

}  // namespace

struct tt_fft_plan_s {
    int nx, ny, nz, nxh;
    AxisPlan ax[3];
    float2 *tw[3];
    float2 *tw_all;
};

extern "C" int tt_fft_plan_create(int32_t nx, int32_t ny, int32_t nz, tt_fft_plan *plan) {
    TT_REQUIRE(plan, "tt_fft_plan_create: null plan pointer");
    TT_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2 && nx <= 4096 && ny <= 4096 && nz <= 4096,
               "tt_fft_plan_create: sizes must be in [2, 4096], got %d x %d x %d", nx, ny, nz);
    TT_REQUIRE(nx <= kBatch && ny <= 65535 && nz <= 65535, "tt_fft_plan_create: size too large");
    tt_fft_plan_s *p = new tt_fft_plan_s();
    p->nx = nx;
    p->ny = ny;
    p->nz = nz;
    p->nxh = nx / 2 + 1;
    const int ns[3] = {nx, ny, nz};
    std::vector<float2> host;
    size_t off[3];
    for (int a = 0; a < 3; ++a) {
        factorize(ns[a], p->ax[a]);
        if (p->ax[a].nrad > kMaxRad) {
            delete p;
            tt::set_error("tt_fft_plan_create: too many factors");
            return TT_ERR_ARG;
        }
        off[a] = host.size();
        for (int m = 0; m < ns[a]; ++m) {
            const double ang = 2.0 * M_PI * (double)m / (double)ns[a];
            host.push_back(make_float2((float)cos(ang), (float)sin(ang)));
        }
    }
    hipError_t e = hipMalloc(&p->tw_all, host.size() * sizeof(float2));
    if (e != hipSuccess) {
        delete p;
        return tt::hip_check(e, "tt_fft_plan_create: hipMalloc");
    }
    e = hipMemcpy(p->tw_all, host.data(), host.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p->tw_all);
        delete p;
        return tt::hip_check(e, "tt_fft_plan_create: hipMemcpy");
    }
    for (int a = 0; a < 3; ++a) p->tw[a] = p->tw_all + off[a];
    *plan = p;
    return TT_OK;
}

extern "C" int tt_fft_plan_destroy(tt_fft_plan plan) {
    if (!plan) return TT_OK;
    hipError_t e = hipFree(plan->tw_all);
    delete plan;
    return tt::hip_check(e, "tt_fft_plan_destroy: hipFree");
}

extern "C" int tt_fft_workspace_bytes(tt_fft_plan plan, size_t *bytes) {
    TT_REQUIRE(plan && bytes, "tt_fft_workspace_bytes: null pointer");
    *bytes = (size_t)plan->nz * plan->ny * plan->nxh * sizeof(float2);
    return TT_OK;
}

namespace {

#define KERNELS_OF(M)                                                                                   \
    (const void *)fft_pass_kernel<M, 0>, (const void *)fft_pass_kernel<M, 128>,                         \
        (const void *)fft_pass_kernel<M, 192>, (const void *)fft_pass_kernel<M, 256>,                   \
        (const void *)fft_pass_kernel<M, 384>, (const void *)fft_pass_kernel<M, 512>,                   \
        (const void *)fft_pass_kernel<M, 1024>

// dynamic LDS above 64 KB must be enabled per kernel
int enable_lds() {
    static int rc = -1;
    if (rc >= 0) return rc;
    // the largest request: n = 4096 (2 pencils per block), rounded up to 1 KB
    // (98 KB + 32 KB of power-spectrum bins, within the 160 KB of a CU)
    const int want = (int)((smem_bytes(4096, pencils_per_block(4096)) + 2 * kMaxBins * sizeof(double) + 1023) /
                           1024 * 1024);
    const void *ks[] = {KERNELS_OF(kGenNative), KERNELS_OF(kGenNoise), KERNELS_OF(kPassY),
                        KERNELS_OF(kPassX),     KERNELS_OF(kFwdX),     KERNELS_OF(kPowZ)};
    for (const void *k : ks) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, want);
        if (e != hipSuccess) return rc = tt::hip_check(e, "hipFuncSetAttribute(max dynamic LDS)");
    }
    return rc = TT_OK;
}

// one pass over pencils of length n: the fixed-length kernel for 128, 192,
// 256, 384, 512 and 1024, the general one otherwise
template <int MODE>
void launch_pass(int n, dim3 grid, size_t sm, hipStream_t st, const PassArgs &a) {
    switch (n) {
        case 128: hipLaunchKernelGGL((fft_pass_kernel<MODE, 128>), grid, dim3(kThreads), sm, st, a); break;
        case 192: hipLaunchKernelGGL((fft_pass_kernel<MODE, 192>), grid, dim3(kThreads), sm, st, a); break;
        case 256: hipLaunchKernelGGL((fft_pass_kernel<MODE, 256>), grid, dim3(kThreads), sm, st, a); break;
        case 384: hipLaunchKernelGGL((fft_pass_kernel<MODE, 384>), grid, dim3(kThreads), sm, st, a); break;
        case 512: hipLaunchKernelGGL((fft_pass_kernel<MODE, 512>), grid, dim3(kThreads), sm, st, a); break;
        case 1024: hipLaunchKernelGGL((fft_pass_kernel<MODE, 1024>), grid, dim3(kThreads), sm, st, a); break;
        default: hipLaunchKernelGGL((fft_pass_kernel<MODE, 0>), grid, dim3(kThreads), sm, st, a); break;
    }
}

int run_passes(tt_fft_plan plan, PassArgs a, bool native, void *workspace, float *out, double *sumsq,
               hipStream_t st) {
    if (int rc = enable_lds()) return rc;
    a.nx = plan->nx;
    a.ny = plan->ny;
    a.nz = plan->nz;
    a.nxh = plan->nxh;
    a.spec = static_cast<float2 *>(workspace);
    a.out = out;
    a.sumsq = sumsq;
    a.scale = (float)(1.0 / ((double)plan->nx * plan->ny * plan->nz));
    // pass Z (with generation)
    a.ap = plan->ax[2];
    a.tw = plan->tw[2];
    a.B = pencils_per_block(plan->nz);
    {
        dim3 grid((unsigned)tt::ceil_div(plan->nxh, a.B), (unsigned)plan->ny);
        const size_t sm = smem_bytes(plan->nz, a.B);
        if (native)
            launch_pass<kGenNative>(plan->nz, grid, sm, st, a);
        else
            launch_pass<kGenNoise>(plan->nz, grid, sm, st, a);
        TT_LAUNCH_CHECK("field pass Z launch");
    }
    // pass Y
    a.ap = plan->ax[1];
    a.tw = plan->tw[1];
    a.B = pencils_per_block(plan->ny);
    {
        dim3 grid((unsigned)tt::ceil_div(plan->nxh, a.B), (unsigned)plan->nz);
        launch_pass<kPassY>(plan->ny, grid, smem_bytes(plan->ny, a.B), st, a);
        TT_LAUNCH_CHECK("field pass Y launch");
    }
    // pass X
    a.ap = plan->ax[0];
    a.tw = plan->tw[0];
    a.B = pencils_per_block(plan->nx);
    {
        dim3 grid((unsigned)tt::ceil_div((int64_t)plan->ny * plan->nz, a.B));
        launch_pass<kPassX>(plan->nx, grid, smem_bytes(plan->nx, a.B), st, a);
        TT_LAUNCH_CHECK("field pass X launch");
    }
    return TT_OK;
}

}  // namespace

extern "C" int tt_field_generate(tt_fft_plan plan, const tt_spectrum *spec, void *workspace, float *out,
                                 double *sumsq, void *stream) {
    TT_REQUIRE(plan && spec && workspace && out, "tt_field_generate: null pointer");
    TT_REQUIRE(spec->aniso > 0, "tt_field_generate: aniso must be > 0");
    TT_REQUIRE(spec->amp >= 0, "tt_field_generate: amp must be >= 0");
    PassArgs a{};
    a.aniso = spec->aniso;
    a.kmin = spec->kmin;
    a.kmax = spec->kmax;
    a.half_beta = (float)(0.5 * spec->beta);
    a.sqrt_amp = (float)sqrt(spec->amp);
    a.key0 = (uint32_t)spec->seed;
    a.key1 = (uint32_t)(spec->seed >> 32);
    return run_passes(plan, a, true, workspace, out, sumsq, tt::as_stream(stream));
}

extern "C" int tt_field_from_noise(tt_fft_plan plan, const float *wr, const float *wi, const float *sqrt_psd,
                                   void *workspace, float *out, double *sumsq, void *stream) {
    TT_REQUIRE(plan && wr && wi && sqrt_psd && workspace && out, "tt_field_from_noise: null pointer");
    PassArgs a{};
    a.wr = wr;
    a.wi = wi;
    a.sp = sqrt_psd;
    return run_passes(plan, a, false, workspace, out, sumsq, tt::as_stream(stream));
}

extern "C" int tt_power_spectrum(tt_fft_plan plan, const float *field, int32_t mode, int32_t nbins, double dx,
                                 double bin_w, void *workspace, double *sums, double *counts, void *stream) {
    TT_REQUIRE(plan && field && workspace && sums && counts, "tt_power_spectrum: null pointer");
    TT_REQUIRE(mode == 0 || mode == 1, "tt_power_spectrum: mode must be 0 (spectrum_3D_scalar) or 1 (compute3Dspectrum)");
    TT_REQUIRE(nbins >= 1 && nbins <= kMaxBins, "tt_power_spectrum: nbins must be in [1, %d]", kMaxBins);
    TT_REQUIRE(dx > 0, "tt_power_spectrum: dx must be > 0");
    TT_REQUIRE(mode == 1 || bin_w > 0, "tt_power_spectrum: bin_w must be > 0");
    if (int rc = enable_lds()) return rc;
    hipStream_t st = tt::as_stream(stream);
    PassArgs a{};
    a.nx = plan->nx;
    a.ny = plan->ny;
    a.nz = plan->nz;
    a.nxh = plan->nxh;
    a.spec = static_cast<float2 *>(workspace);
    a.in = field;
    a.mode = mode;
    a.nbins = nbins;
    // numpy fftfreq(M, d): integer index times 1.0 / (M * d)
    a.fx_step = 1.0 / ((double)plan->nx * dx);
    a.fy_step = 1.0 / ((double)plan->ny * dx);
    a.fz_step = 1.0 / ((double)plan->nz * dx);
    a.bin_w = bin_w;
    const double nt = (double)plan->nx * plan->ny * plan->nz;
    a.pscale = 1.0 / (nt * nt);
    a.hsum = sums;
    a.hcnt = counts;
    // pass X: real rows -> half spectrum
    a.ap = plan->ax[0];
    a.tw = plan->tw[0];
    a.B = pencils_per_block(plan->nx);
    launch_pass<kFwdX>(plan->nx, dim3((unsigned)tt::ceil_div((int64_t)plan->ny * plan->nz, a.B)),
                       smem_bytes(plan->nx, a.B), st, a);
    TT_LAUNCH_CHECK("power spectrum pass X launch");
    // pass Y
    a.ap = plan->ax[1];
    a.tw = plan->tw[1];
    a.B = pencils_per_block(plan->ny);
    launch_pass<kPassY>(plan->ny, dim3((unsigned)tt::ceil_div(plan->nxh, a.B), (unsigned)plan->nz),
                        smem_bytes(plan->ny, a.B), st, a);
    TT_LAUNCH_CHECK("power spectrum pass Y launch");
    // pass Z with |F|^2 binning (sums and counts are ADDED to)
    a.ap = plan->ax[2];
    a.tw = plan->tw[2];
    a.B = pencils_per_block(plan->nz);
    const size_t sm = smem_bytes(plan->nz, a.B) + 2 * (size_t)nbins * sizeof(double);
    launch_pass<kPowZ>(plan->nz, dim3((unsigned)tt::ceil_div(plan->nxh, a.B), (unsigned)plan->ny), sm, st, a);
    TT_LAUNCH_CHECK("power spectrum pass Z launch");
    return TT_OK;
}
