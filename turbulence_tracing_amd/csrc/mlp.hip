// Surrogate MLP kernels: BF16 MFMA GEMM with fused bias/GELU epilogue
// (forward and both backward products), GELU backward, bias gradient, loss
// gradient and Adam.  No reference counterpart: the reference has no learned
// component; these serve the learned correction of the cheap path estimate
// (tt_path_features) against the full RK4 trace.
//
// GEMM: C[m][n] = epi( sum_k A(m,k) B(k,n) + bias[n] ), A and B bf16 with
// arbitrary element strides, fp32 accumulation on v_mfma_f32_16x16x32_bf16.
// blockIdx.x tiles M (the long ray dimension), blockIdx.y tiles N.
// A block computes a 64 x 64 tile with 4 waves (2 x 2 of 32 x 32, i.e. 2 x 2
// MFMA tiles each); K advances 32 at a time through LDS tiles stored
// K-contiguous so every lane reads its 8-element MFMA fragment as one 16-byte
// LDS load.  Fragment layout (16x16x32): lane l holds A[l%16][8(l/16)+j],
// B[8(l/16)+j][l%16], and D[4(l/16)+i][l%16].  blockIdx.z splits K; with
// split-K the partial tiles are added atomically into fp32 C.
#include <hip/hip_bf16.h>

#include "tt_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));


namespace {

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int PADK = 8;  // LDS row padding (bf16 elements) to spread banks

// GELU, tanh form: 0.5 x (1 + tanh(u)), u = sqrt(2/pi) (x + 0.044715 x^3),
// evaluated as x * s with s = 0.5 (1 + tanh(u)) = 1 / (1 + exp(-2u)) and
// exp(-2u) = 2^(x (a + b x^2)), a = -2 sqrt(2/pi) log2(e), b = 0.044715 a.
// Saturates correctly: s -> 0 (exp2 -> inf) and s -> 1 (exp2 -> 0).
__device__ __forceinline__ float gelu_s(float x) {
    const float e = __builtin_amdgcn_exp2f(x * fmaf(-0.1029432395800235f, x * x, -2.302208198144325f));
    return __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu_f(float x) { return x * gelu_s(x); }
// d/dx [x s(2u)] = s + x 2 s (1 - s) u', u' = sqrt(2/pi) (1 + 3 * 0.044715 x^2)
__device__ __forceinline__ float gelu_grad(float x) {
    const float s = gelu_s(x);
    const float du = 0.7978845608028654f * fmaf(0.134145f, x * x, 1.0f);
    return fmaf(2.0f * x * s * (1.0f - s), du, s);
}

// Fill a [64 rows][32 k] LDS tile of operand X(row, k) = X[row*srow + k*sk].
// mode 1: contiguous along k (8 consecutive k per thread); mode 2: contiguous
// along rows (8 consecutive rows per thread); mode 0: element by element.
// Vector loads are used only when the whole 8-element run is in range.
__device__ __forceinline__ void fill_tile(const __bf16 *X, int64_t srow, int64_t sk, int rows, int r0, int k0, int kend,
                                          int mode, __bf16 (*T)[BK + PADK], int tid) {
    typedef __bf16 bf16x8_u __attribute__((ext_vector_type(8)));
    if (mode == 2) {
        const int k = tid >> 3, r8 = (tid & 7) * 8;
        const int kk = k0 + k, rr = r0 + r8;
        if (kk < kend && rr + 8 <= rows) {
            const bf16x8_u v = *reinterpret_cast<const bf16x8_u *>(X + (int64_t)rr * srow + (int64_t)kk * sk);
#pragma unroll
            for (int j = 0; j < 8; ++j) T[r8 + j][k] = v[j];
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                T[r8 + j][k] = (kk < kend && rr + j < rows) ? X[(int64_t)(rr + j) * srow + (int64_t)kk * sk]
                                                           : (__bf16)0.0f;
        }
        return;
    }
    const int r = tid >> 2, c8 = (tid & 3) * 8;
    const int rr = r0 + r, kk = k0 + c8;
    if (mode == 1 && rr < rows && kk + 8 <= kend) {
        *reinterpret_cast<bf16x8_u *>(&T[r][c8]) =
            *reinterpret_cast<const bf16x8_u *>(X + (int64_t)rr * srow + (int64_t)kk * sk);
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j)
            T[r][c8 + j] = (rr < rows && kk + j < kend) ? X[(int64_t)rr * srow + (int64_t)(kk + j) * sk]
                                                        : (__bf16)0.0f;
    }
}

__device__ __forceinline__ int vec_mode(const void *p, int64_t srow, int64_t sk) {
    const bool al = ((uintptr_t)p & 15) == 0;
    if (al && sk == 1 && srow % 8 == 0) return 1;
    if (al && srow == 1 && sk % 8 == 0) return 2;
    return 0;
}

__global__ __launch_bounds__(256) void gemm_kernel(tt_gemm_args g) {
    __shared__ __bf16 As[BM][BK + PADK];  // [m][k]
    __shared__ __bf16 Bs[BN][BK + PADK];  // [n][k]
    const __bf16 *A = static_cast<const __bf16 *>(g.A);
    const __bf16 *B = static_cast<const __bf16 *>(g.B);
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int a_mode = vec_mode(g.A, g.sam, g.sak), b_mode = vec_mode(g.B, g.sbn, g.sbk);
    int kbeg = 0, kend = g.K;
    if (g.ksplit > 0) {
        kbeg = blockIdx.z * g.ksplit;
        kend = min(g.K, kbeg + g.ksplit);
    }
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        // fill the tiles: 64 x 32 elements each, 8 per thread, as one 16-byte
        // load when the operand is contiguous along k or along m / n
        fill_tile(A, g.sam, g.sak, g.M, m0, k0, kend, a_mode, As, tid);
        fill_tile(B, g.sbn, g.sbk, g.N, n0, k0, kend, b_mode, Bs, tid);
        __syncthreads();
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) af[i] = *reinterpret_cast<const bf16x8 *>(&As[wm + 16 * i + fr][fk]);
#pragma unroll
        for (int j = 0; j < 2; ++j) bfr[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + 16 * j + fr][fk]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    // epilogue: lane holds D[4(l/16)+e][l%16] of each 16 x 16 tile
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + e;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][e];
                const int64_t o = (int64_t)m * g.ldc + n;
                if (g.ksplit > 0) {
                    atomicAdd(&g.Cf[o], v);
                    continue;
                }
                if (g.bias) v += g.bias[n];
                if (g.Zpre) g.Zpre[o] = v;
                if (g.gelu) v = gelu_f(v);
                if (g.Cf) g.Cf[o] = v;
                if (g.Cb) static_cast<__bf16 *>(g.Cb)[o] = (__bf16)v;
            }
}

// dZ = dH * gelu'(Zpre) -> bf16 (elementwise; dH fp32 or bf16)
__global__ void gelu_bwd_kernel(const float *dHf, const __bf16 *dHb, const float *Z, __bf16 *dZ, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = dHf ? dHf[i] : (float)dHb[i];
    dZ[i] = (__bf16)(d * gelu_grad(Z[i]));
}

// db[c] += sum_r dZ[r][c] for a block of rows (atomics across blocks)
__global__ void colsum_kernel(const __bf16 *dZ, int64_t rows, int cols, int64_t ld, float *db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    const int64_t r0 = (int64_t)blockIdx.y * 256;
    const int64_t r1 = min(rows, r0 + 256);
    float s = 0.f;
    for (int64_t r = r0; r < r1; ++r) s += (float)dZ[r * ld + c];
    atomicAdd(&db[c], s);
}

// Loss over the first n_out columns of Y (fp32, ld ldy) against targets T
// (fp32, ld ldt): columns [0, n_reg) are weighted squared error, columns
// [n_reg, n_reg + n_err) predict |T - Y| of column group err_of[j] (L1, the
// regression output treated as a constant).  dY (bf16, ld ldd) receives the
// gradient of loss / rows; loss_out (2 doubles) accumulates
// sum(w (Y - T)^2) and sum |E - |T - Y||.
struct LossArgs {
    const float *Y;
    int64_t ldy;
    const float *T;
    int64_t ldt;
    __bf16 *dY;
    int64_t ldd;
    int64_t rows;
    int n_reg, n_err, n_pad;
    float w[8];
    int err_lo[4], err_hi[4];  // error head j predicts the max |T - Y| over columns [lo, hi)
    double *loss_out;
    float tau;
};

__global__ void loss_kernel(LossArgs a) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s2 = 0.0, s1 = 0.0;
    if (r < a.rows) {
        const float *y = a.Y + r * a.ldy;
        const float *t = a.T + r * a.ldt;
        __bf16 *d = a.dY + r * a.ldd;
        const float inv = 1.0f / (float)a.rows;
        for (int j = 0; j < a.n_reg; ++j) {
            const float e = y[j] - t[j];
            s2 += (double)(a.w[j] * e * e);
            d[j] = (__bf16)(2.0f * a.w[j] * e * inv);
        }
        // pinball loss at quantile tau: the head learns the tau-quantile of
        // the corrected error max |T - Y| over its columns
        for (int h = 0; h < a.n_err; ++h) {
            float target = 0.f;
            for (int j = a.err_lo[h]; j < a.err_hi[h]; ++j) target = fmaxf(target, fabsf(t[j] - y[j]));
            const float e = y[a.n_reg + h] - target;
            s1 += (double)(e > 0.f ? (1.0f - a.tau) * e : -a.tau * e);
            d[a.n_reg + h] = (__bf16)((e > 0.f ? 1.0f - a.tau : (e < 0.f ? -a.tau : 0.0f)) * inv);
        }
        for (int j = a.n_reg + a.n_err; j < a.n_pad; ++j) d[j] = (__bf16)0.0f;
    }
    // wave reduction, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        s2 += __shfl_down(s2, off, 64);
        s1 += __shfl_down(s1, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&a.loss_out[0], s2);
        atomicAdd(&a.loss_out[1], s1);
    }
}

// Adam (Kingma & Ba 2015) on fp32 master weights, refreshing the bf16 copy
__global__ void adam_kernel(float *w, const float *g, float *m, float *v, __bf16 *wb, int64_t n, float lr, float b1,
                            float b2, float eps, float c1, float c2, float wd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = (__bf16)wi;
}

// (x - mu) / sigma -> bf16, row-major [rows][cols] into [rows][ldo] (zero pad)
__global__ void standardize_kernel(const float *x, int64_t srow, int64_t scol, int64_t rows, int cols,
                                   const float *mu, const float *isd, __bf16 *out, int64_t ldo) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = i / ldo;
    const int c = (int)(i - r * ldo);
    if (r >= rows) return;
    float v = 0.f;
    if (c < cols) v = (x[r * srow + c * scol] - mu[c]) * isd[c];
    out[i] = (__bf16)v;
}

// ------------------------------------------------------------ inference
// Fused forward pass of the surrogate network in[<=64] -> 256 -> 256 -> 256
// -> out[<=16] (GELU on the hidden layers), input standardisation included.
// Computed transposed, H'^T = W H^T: the weights are the MFMA A operand and
// the activations of 16 rays the B operand, so a layer's D tiles (lane l:
// feature 4(l/16)+i of a 16-feature tile, ray l%16) become the next layer's
// B operand in registers with no data movement: the 8 bf16 of lane l for
// k-chunk c are features 32c + 16(j/4) + 4(l/16) + j%4, and the next layer's
// weight columns are packed in that order (kperm) on the host.  The first
// layer reads features 32c + 8(l/16) + j directly.
//
// Packed weights: per layer, fragments (tile nt, chunk c) in order
// nt * KC + c, each 64 lanes x 8 bf16 (lane l: W[16nt + l%16][kperm(c, l/16,
// j)]), zero-padded.  Layer shapes (NT x KC): 16 x 2, 16 x 8, 16 x 8, 1 x 8.
// A block of 4 waves takes 128 rays (32 per wave, two 16-ray B tiles).  The
// 296 weight fragments (1 KB each, all four layers in one stream) pass
// through two 16 KB LDS buffers in chunks of 16 fragments: the copy of chunk
// c + 1 (global -> LDS direct, no registers) is in flight while chunk c is
// computed, and the waves of a block meet once per chunk at an LDS-only
// barrier, so the copy's latency hides behind a chunk of MFMAs.  Each
// fragment is read once from L2 per block and once from LDS per wave for two
// MFMAs.  The accumulators start from the bias (kept in LDS), so the epilogue
// is GELU and the bf16 conversion only.  160 VGPRs: three blocks per CU.
constexpr int kInferTiles = 2;                // 16-ray tiles per wave
constexpr int kInferRays = 64 * kInferTiles;  // rays per block of 4 waves
constexpr int kFragsTotal = 16 * 2 + 16 * 8 + 16 * 8 + 8;  // 296 fragments of 1 KB
constexpr int kPipeG = 16;                    // fragments per chunk

struct InferArgs {
    const float *feat;    // element (ray r, feature f) at feat[r * srow + f * scol]
    int64_t srow, scol, n;
    int nfeat, nout;
    const float *mu, *isd;
    const __bf16 *w;
    const float *bias;  // 256 + 256 + 256 + 16
    float *Y;
    int64_t ldy;
};

typedef __attribute__((address_space(1))) void global_void;
typedef __attribute__((address_space(3))) void lds_void;

// issue the copy of fragment chunk `chunk` (G fragments, fewer at the end)
// into LDS buffer `buf`: lane l of wave v moves 16 bytes per instruction to
// buf + (256 i + 64 v + l) * 16
template <int G>
__device__ __forceinline__ void copy_chunk(const __bf16 *w, int chunk, __bf16 *buf) {
    const int nfr = kFragsTotal - chunk * G < G ? kFragsTotal - chunk * G : G;
    if (nfr <= 0) return;
    const uint4 *src = reinterpret_cast<const uint4 *>(w) + (size_t)chunk * G * 64;
    uint4 *dst = reinterpret_cast<uint4 *>(buf);
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < G * 64 / 256; ++i)
        if (i * 256 < nfr * 64)
            __builtin_amdgcn_global_load_lds((global_void *)(src + i * 256 + threadIdx.x),
                                             (lds_void *)(dst + i * 256 + wave * 64), 16, 0, 0);
}

// LDS-only workgroup barrier: the fences order LDS operations only, so the
// barrier does not wait for this wave's global loads
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The two buffers are separate LDS arrays, so the compiler can tell a copy
// into one from the reads of the other and does not wait for the copy first.
template <int G>
struct Pipe {
    const __bf16 *w;
    __bf16 *b0, *b1;
    __device__ __forceinline__ __bf16 *buf(int c) const { return (c & 1) ? b1 : b0; }
    // chunk c is about to be computed: start the copy of chunk c + 1
    __device__ __forceinline__ void ahead(int c) const { copy_chunk<G>(w, c + 1, buf(c + 1)); }
    // chunk c is computed: this thread's copies of chunk c + 1 have landed
    // (vmcnt 0), then every wave has finished with chunk c and copied c + 1
    __device__ __forceinline__ void done(int) const {
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt / lgkmcnt not waited on
        lds_barrier();
    }
};

// one layer of NT output tiles x KC input chunks from the fragment stream,
// chunks C0 .. C0 + NT*KC/G - 1
template <int KC, int T, int NT, int C0, int G, bool GELU>
__device__ __forceinline__ void pipe_layer(const Pipe<G> &P, const float *bias, const bf16x8 (&x)[KC][T],
                                           bf16x8 (&y)[8][T], f32x4 (&out)[T]) {
    constexpr int NG = G / KC;  // output tiles per chunk
    const int lane = threadIdx.x & 63, g = lane >> 4;
#pragma unroll
    for (int grp = 0; grp < (NT + NG - 1) / NG; ++grp) {
        const int c = C0 + grp;
        P.ahead(c);
        const __bf16 *lb = P.buf(c);
#pragma unroll
        for (int tl = 0; tl < NG; ++tl) {
            const int nt = grp * NG + tl;
            if (nt >= NT) break;
            f32x4 acc[T];
            if (GELU) {  // bias from LDS: a global load here would wait on the copies in flight
                const float4 bb = *reinterpret_cast<const float4 *>(bias + 16 * nt + 4 * g);
#pragma unroll
                for (int t = 0; t < T; ++t) acc[t] = f32x4{bb.x, bb.y, bb.z, bb.w};
            } else {
#pragma unroll
                for (int t = 0; t < T; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
            }
            // raised wave priority while this wave issues its MFMA chain, so the
            // SIMD's other waves' epilogues fill the gaps rather than delay it
            __builtin_amdgcn_s_setprio(1);
            // the fragments in halves of up to four: a half's LDS reads are all
            // issued before its MFMAs (sched_barrier keeps the scheduler from
            // sinking each read to its use, which left two reads per wait)
            constexpr int HALF = KC < 4 ? KC : 4;
#pragma unroll
            for (int c0 = 0; c0 < KC; c0 += HALF) {
                bf16x8 af[HALF];
#pragma unroll
                for (int j = 0; j < HALF; ++j)
                    af[j] = *reinterpret_cast<const bf16x8 *>(lb + ((tl * KC + c0 + j) * 64 + lane) * 8);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < HALF; ++j)
#pragma unroll
                    for (int t = 0; t < T; ++t)
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[j], x[c0 + j][t], acc[t], 0, 0, 0);
            }
            __builtin_amdgcn_s_setprio(0);
            if (GELU) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int t = 0; t < T; ++t) y[nt >> 1][t][4 * (nt & 1) + i] = (__bf16)gelu_f(acc[t][i]);
            } else {
#pragma unroll
                for (int t = 0; t < T; ++t) out[t] = acc[t];
            }
        }
        P.done(c);
    }
}

template <int T, int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void mlp_infer_kernel(InferArgs a) {
    __shared__ __attribute__((aligned(16))) __bf16 lds0[G * 512], lds1[G * 512];
    __shared__ __attribute__((aligned(16))) float lbias[768];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    const Pipe<G> P{a.w, lds0, lds1};
    copy_chunk<G>(a.w, 0, P.buf(0));
    for (int j = threadIdx.x; j < 768; j += 256) lbias[j] = a.bias[j];
    const int64_t ray0 = (int64_t)blockIdx.x * (64 * T) + wave * (16 * T);
    bf16x8 x0[2][T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int64_t ray = ray0 + 16 * t + r;
        const float *row = a.feat + ray * a.srow;
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int f = 32 * c + 8 * g + j;
                float v = 0.f;
                if (ray < a.n && f < a.nfeat) v = (row[f * a.scol] - a.mu[f]) * a.isd[f];
                x0[c][t][j] = (__bf16)v;
            }
    }
    __syncthreads();  // features, biases and chunk 0 are in
    bf16x8 h1[8][T], h2[8][T];
    f32x4 acc[T];
    constexpr int C1 = 0, C2 = C1 + 16 * 2 / G, C3 = C2 + 16 * 8 / G, C4 = C3 + 16 * 8 / G;
    pipe_layer<2, T, 16, C1, G, true>(P, lbias, x0, h1, acc);
    pipe_layer<8, T, 16, C2, G, true>(P, lbias + 256, h1, h2, acc);
    pipe_layer<8, T, 16, C3, G, true>(P, lbias + 512, h2, h1, acc);
    pipe_layer<8, T, 1, C4, G, false>(P, lbias, h1, h2, acc);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int o = 4 * g + i;
        if (o >= a.nout) continue;
        const float b = a.bias[768 + o];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int64_t ra = ray0 + 16 * t + r;
            if (ra < a.n) a.Y[ra * a.ldy + o] = acc[t][i] + b;
        }
    }
}

}  // namespace

extern "C" int tt_gemm_bf16(const tt_gemm_args *g, void *stream) {
    TT_REQUIRE(g && g->A && g->B, "tt_gemm_bf16: null operand");
    TT_REQUIRE(g->M >= 0 && g->N >= 0 && g->K >= 0, "tt_gemm_bf16: negative size");
    TT_REQUIRE(g->ksplit == 0 || (g->Cf && !g->Cb && !g->Zpre && !g->gelu && !g->bias),
               "tt_gemm_bf16: split-K accumulates raw products into fp32 Cf only");
    TT_REQUIRE(g->ksplit == 0 || g->ksplit % BK == 0, "tt_gemm_bf16: ksplit must be a multiple of %d", BK);
    TT_REQUIRE(g->Cf || g->Cb, "tt_gemm_bf16: no output");
    TT_REQUIRE(g->ldc >= g->N, "tt_gemm_bf16: ldc < N");
    if (g->M == 0 || g->N == 0) return TT_OK;
    const int64_t zs = g->ksplit > 0 ? tt::ceil_div(g->K, g->ksplit) : 1;
    TT_REQUIRE(zs <= 65535, "tt_gemm_bf16: too many K chunks");
    dim3 grid((unsigned)tt::ceil_div(g->M, BM), (unsigned)tt::ceil_div(g->N, BN), (unsigned)zs);
    TT_REQUIRE(grid.y <= 65535u, "tt_gemm_bf16: N too large");
    hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, tt::as_stream(stream), *g);
    TT_LAUNCH_CHECK("tt_gemm_bf16 launch");
    return TT_OK;
}

extern "C" int tt_gelu_backward(const float *dH_f32, const void *dH_bf16, const float *Zpre, void *dZ_bf16, int64_t n,
                                void *stream) {
    TT_REQUIRE((dH_f32 || dH_bf16) && Zpre && dZ_bf16 && n >= 0, "tt_gelu_backward: bad arguments");
    if (n == 0) return TT_OK;
    hipLaunchKernelGGL(gelu_bwd_kernel, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0, tt::as_stream(stream),
                       dH_f32, static_cast<const __bf16 *>(dH_bf16), Zpre, static_cast<__bf16 *>(dZ_bf16), n);
    TT_LAUNCH_CHECK("tt_gelu_backward launch");
    return TT_OK;
}

extern "C" int tt_bias_grad(const void *dZ_bf16, int64_t rows, int32_t cols, int64_t ld, float *db, void *stream) {
    TT_REQUIRE(dZ_bf16 && db && rows >= 0 && cols > 0 && ld >= cols, "tt_bias_grad: bad arguments");
    if (rows == 0) return TT_OK;
    dim3 grid((unsigned)tt::ceil_div(cols, 64), (unsigned)tt::ceil_div(rows, 256));
    TT_REQUIRE(grid.y <= 65535u * 64u, "tt_bias_grad: too many rows");
    hipLaunchKernelGGL(colsum_kernel, grid, dim3(64), 0, tt::as_stream(stream), static_cast<const __bf16 *>(dZ_bf16),
                       rows, cols, ld, db);
    TT_LAUNCH_CHECK("tt_bias_grad launch");
    return TT_OK;
}

extern "C" int tt_surrogate_loss(const float *Y, int64_t ldy, const float *T, int64_t ldt, void *dY_bf16,
                                 int64_t ldd, int64_t rows, int32_t n_reg, const float *w_reg, int32_t n_err,
                                 const int32_t *err_lo, const int32_t *err_hi, double tau, int32_t n_pad, double *loss_out,
                                 void *stream) {
    TT_REQUIRE(Y && T && dY_bf16 && loss_out && w_reg, "tt_surrogate_loss: null pointer");
    TT_REQUIRE(tau > 0.0 && tau < 1.0, "tt_surrogate_loss: tau must be in (0, 1)");
    TT_REQUIRE(n_reg >= 1 && n_reg <= 8 && n_err >= 0 && n_err <= 4 && n_pad >= n_reg + n_err && n_pad <= ldd,
               "tt_surrogate_loss: bad column counts");
    TT_REQUIRE(ldy >= n_reg + n_err && ldt >= n_reg, "tt_surrogate_loss: bad leading dimensions");
    LossArgs a{};
    a.Y = Y;
    a.ldy = ldy;
    a.T = T;
    a.ldt = ldt;
    a.dY = static_cast<__bf16 *>(dY_bf16);
    a.ldd = ldd;
    a.rows = rows;
    a.n_reg = n_reg;
    a.n_err = n_err;
    a.n_pad = n_pad;
    for (int j = 0; j < n_reg; ++j) a.w[j] = w_reg[j];
    for (int h = 0; h < n_err; ++h) {
        TT_REQUIRE(err_lo && err_hi && err_lo[h] >= 0 && err_hi[h] <= n_reg && err_lo[h] < err_hi[h],
                   "tt_surrogate_loss: bad error-head column range");
        a.err_lo[h] = err_lo[h];
        a.err_hi[h] = err_hi[h];
    }
    a.loss_out = loss_out;
    a.tau = (float)tau;
    if (rows == 0) return TT_OK;
    hipLaunchKernelGGL(loss_kernel, dim3((unsigned)tt::ceil_div(rows, 256)), dim3(256), 0, tt::as_stream(stream), a);
    TT_LAUNCH_CHECK("tt_surrogate_loss launch");
    return TT_OK;
}

extern "C" int tt_adam(float *w, const float *g, float *m, float *v, void *w_bf16, int64_t n, double lr, double b1,
                       double b2, double eps, int64_t step, double weight_decay, void *stream) {
    TT_REQUIRE(w && g && m && v && n >= 0 && step >= 1, "tt_adam: bad arguments");
    if (n == 0) return TT_OK;
    const float c1 = (float)(1.0 - pow(b1, (double)step)), c2 = (float)(1.0 - pow(b2, (double)step));
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0, tt::as_stream(stream), w, g, m,
                       v, static_cast<__bf16 *>(w_bf16), n, (float)lr, (float)b1, (float)b2, (float)eps, c1, c2,
                       (float)weight_decay);
    TT_LAUNCH_CHECK("tt_adam launch");
    return TT_OK;
}

extern "C" int tt_standardize_bf16(const float *x, int64_t srow, int64_t scol, int64_t rows, int32_t cols,
                                   const float *mu, const float *inv_sd, void *out_bf16, int64_t ldo, void *stream) {
    TT_REQUIRE(x && mu && inv_sd && out_bf16 && rows >= 0 && cols > 0 && ldo >= cols && srow >= 1 && scol >= 1,
               "tt_standardize_bf16: bad arguments");
    if (rows == 0) return TT_OK;
    const int64_t n = rows * ldo;
    hipLaunchKernelGGL(standardize_kernel, dim3((unsigned)tt::ceil_div(n, 256)), dim3(256), 0, tt::as_stream(stream),
                       x, srow, scol, rows, cols, mu, inv_sd, static_cast<__bf16 *>(out_bf16), ldo);
    TT_LAUNCH_CHECK("tt_standardize_bf16 launch");
    return TT_OK;
}

extern "C" int tt_mlp_infer(const float *feat, int64_t srow, int64_t scol, int64_t n, int32_t nfeat, const float *mu,
                            const float *inv_sd, const void *w_packed, const float *bias_packed, float *Y, int64_t ldy,
                            int32_t nout, void *stream) {
    TT_REQUIRE(n >= 0, "tt_mlp_infer: n < 0");
    if (n == 0) return TT_OK;  // an empty ray set: nothing is read or written
    TT_REQUIRE(feat && mu && inv_sd && w_packed && bias_packed && Y, "tt_mlp_infer: null pointer");
    TT_REQUIRE(nfeat > 0 && nfeat <= 64, "tt_mlp_infer: need 0 < nfeat <= 64");
    TT_REQUIRE(srow >= 1 && scol >= 1, "tt_mlp_infer: strides must be >= 1");
    TT_REQUIRE(nout > 0 && nout <= 16 && ldy >= nout, "tt_mlp_infer: need 0 < nout <= 16 and ldy >= nout");
    TT_REQUIRE(((uintptr_t)w_packed & 15) == 0 && ((uintptr_t)bias_packed & 15) == 0,
               "tt_mlp_infer: packed weights and biases must be 16-byte aligned");
    InferArgs a;
    a.feat = feat;
    a.srow = srow;
    a.scol = scol;
    a.n = n;
    a.nfeat = nfeat;
    a.nout = nout;
    a.mu = mu;
    a.isd = inv_sd;
    a.w = static_cast<const __bf16 *>(w_packed);
    a.bias = bias_packed;
    a.Y = Y;
    a.ldy = ldy;
    const int64_t blocks = tt::ceil_div(n, kInferRays);
    TT_REQUIRE(blocks <= 0x7fffffff, "tt_mlp_infer: too many rays");
    hipLaunchKernelGGL((mlp_infer_kernel<kInferTiles, kPipeG>), dim3((unsigned)blocks), dim3(256), 0,
                       tt::as_stream(stream), a);
    TT_LAUNCH_CHECK("tt_mlp_infer launch");
    return TT_OK;
}
