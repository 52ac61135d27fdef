// Kernels of the learned warm start: a 3-D convolutional encoder-decoder trained on steady velocity fields
// that the single-phase solver produces (no reference counterpart; the reference starts every run at rest).
//
// Tensors are channels-last, x[n][z][y][x][c] in BF16; channel counts are multiples of 8 so a lane loads 8
// consecutive channels (16 bytes) at once.  A 3x3x3 convolution (stride 1, zero padding 1) is an implicit GEMM
// out[p][co] = sum_k A[p][k] B[k][co] with k = off * cin + ci (off = 9 (dz+1) + 3 (dy+1) + (dx+1)), run on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation).  Weights are stored per output channel as
// rows of kpad = roundup(27 cin, 32) BF16 values (zero beyond 27 cin); the FP32 master copy (rows of 27 cin)
// lives with the optimiser.  The input gradient of a convolution is the same kernel run on the output gradient with the
// kernel flipped and its channel roles swapped (mflbm_nn_weights_to_bf16 writes both forms).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "mflbm.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }
__device__ __forceinline__ __bf16 f2bf(float v) { return (__bf16)v; }

struct Shape {
    int n, d, h, w;
    __host__ __device__ __forceinline__ long count() const { return (long)n * d * h * w; }
};

// Matrix-core fragment layout of v_mfma_f32_16x16x32_bf16 for lane l (64 lanes): A row l % 16 with k
// 8 (l / 16) .. +7; B column l % 16 with the same k; D holds rows 4 (l / 16) + t, t = 0..3, of column l % 16.
__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// 16 x 16 x 32 product of row-major A (16 x 32) and column-major B (columns of 32): the layout check.
__global__ void k_mfma_probe(const __bf16* __restrict__ a, const __bf16* __restrict__ b, float* __restrict__ d) {
    const int l = threadIdx.x;
    const int r = l & 15, q = l >> 4;
    bf16x8 fa, fb;
    for (int j = 0; j < 8; ++j) {
        fa[j] = a[r * 32 + q * 8 + j];
        fb[j] = b[r * 32 + q * 8 + j];
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(fa, fb, acc);
    for (int t = 0; t < 4; ++t) d[(q * 4 + t) * 16 + r] = acc[t];
}

// One wave computes 32 positions x 32 output channels; 4 waves per block along positions.
template <int RELU>
__global__ void __launch_bounds__(NT) k_conv_fwd(const __bf16* __restrict__ x, const __bf16* __restrict__ wt,
                                                 const float* __restrict__ bias, __bf16* __restrict__ y, Shape S,
                                                 int cin, int cout, int kpad) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long M = S.count();
    const long p0 = ((long)blockIdx.x * 4 + wv) * 32;
    const int co0 = blockIdx.y * 32;
    if (p0 >= M) return;
    const int r = lane & 15, q = lane >> 4;
    int pz[2], py[2], px[2], pn[2];
    bool pv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        long p = p0 + h * 16 + r;
        pv[h] = p < M;
        if (!pv[h]) p = 0;
        px[h] = (int)(p % S.w);
        p /= S.w;
        py[h] = (int)(p % S.h);
        p /= S.h;
        pz[h] = (int)(p % S.d);
        pn[h] = (int)(p / S.d);
    }
    const int K = 27 * cin;
    f32x4 acc[2][2];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int g = 0; g < 2; ++g) acc[h][g] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero = {};
    for (int k0 = 0; k0 < kpad; k0 += 32) {
        const int k = k0 + q * 8;
        bf16x8 a[2] = {zero, zero}, b[2] = {zero, zero};
        if (k < K) {
            const int off = k / cin, ci = k - off * cin;
            const int dz = off / 9 - 1, dy = (off / 3) % 3 - 1, dx = off % 3 - 1;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int zz = pz[h] + dz, yy = py[h] + dy, xx = px[h] + dx;
                if (pv[h] && zz >= 0 && zz < S.d && yy >= 0 && yy < S.h && xx >= 0 && xx < S.w)
                    a[h] = *(const bf16x8*)(x + ((((long)pn[h] * S.d + zz) * S.h + yy) * S.w + xx) * cin + ci);
            }
        }
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            const int co = co0 + g * 16 + r;
            if (co < cout) b[g] = *(const bf16x8*)(wt + (long)co * kpad + k);   // zero padded beyond K
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int g = 0; g < 2; ++g) acc[h][g] = mfma(a[h], b[g], acc[h][g]);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            const int co = co0 + g * 16 + r;
            if (co >= cout) continue;
            const float bb = bias ? bias[co] : 0.f;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const long p = p0 + h * 16 + q * 4 + t;
                if (p >= M) continue;
                float v = acc[h][g][t] + bb;
                if (RELU && v < 0.f) v = 0.f;
                y[p * cout + co] = f2bf(v);
            }
        }
}

// Weight gradient: dw[co][k] += sum_p dy[p][co] * x[p + off(k)][ci(k)] (rows of 27 cin, the layout of the FP32
// master weights), the positions split into chunks of `chunk` (a multiple of 32) over grid z.  A work-group computes
// 32 output channels x 128 k (one wave per 32 k) and walks its chunk 32 positions at a time: the dy tile (32
// positions x 32 channels) and the im2col tile of x (32 positions x 128 k) are loaded with 16-byte accesses (8
// channels of one position) and stored transposed in LDS, two buffers in turn (the next tile's loads are in flight
// while the current one is multiplied), and each wave reads its fragments from there.  The fragments, and so the
// sums, are those of a lane-by-lane gather: lane l takes 8 consecutive positions for one channel (A) or one k (B).
// The chunk sums go to dw with FP32 atomics, or (ORDERED) are stored to the chunk's own slice dw + z * cout * K,
// every element written by exactly one lane, for a fixed-order sum after.
constexpr int DW_KT = 128;          // k per work-group
constexpr int DW_LD = 40;           // LDS row length (32 positions + 8: 80-byte rows keep 16-byte fragment reads aligned)
template <bool ORDERED>
__global__ void __launch_bounds__(NT) k_conv_dw(const __bf16* __restrict__ x, const __bf16* __restrict__ dy,
                                                float* __restrict__ dw, Shape S, int cin, int cout, long chunk) {
    __shared__ __attribute__((aligned(16))) __bf16 As[2][32][DW_LD];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][DW_KT][DW_LD];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int co0 = blockIdx.x * 32;
    const int kb0 = blockIdx.y * DW_KT;
    const int K = 27 * cin;
    const long M = S.count();
    const long pbeg = (long)blockIdx.z * chunk;
    const long pend = pbeg + chunk < M ? pbeg + chunk : M;
    const int ntile = (int)((pend - pbeg + 31) / 32);
    const bf16x8 zero = {};
    // this thread's loads of a tile: one dy chunk (threads < 128: position tid / 4, channels 8 (tid % 4) ..) and two
    // x chunks (item u = tid + 256 v: position u / 16, k 8 (u % 16) .. of the work-group's 128)
    auto load = [&](long pk, bf16x8& ra, bf16x8 (&rb)[2]) {
        ra = zero;
        if (tid < 128) {
            const long p = pk + (tid >> 2);
            const int co = co0 + 8 * (tid & 3);
            if (p < pend && co < cout) ra = *(const bf16x8*)(dy + p * cout + co);
        }
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int u = tid + NT * v;
            const long p = pk + (u >> 4);
            const int k = kb0 + 8 * (u & 15);
            rb[v] = zero;
            if (p < pend && k < K) {
                const int off = k / cin, ci = k - off * cin;
                int t = (int)p;   // (positions are below 2^31: 32-bit decode)
                const int xq = t % S.w;
                t /= S.w;
                const int yq = t % S.h;
                t /= S.h;
                const int zq = t % S.d;
                const int nq = t / S.d;
                const int zz = zq + off / 9 - 1, yy = yq + (off / 3) % 3 - 1, xx = xq + off % 3 - 1;
                if (zz >= 0 && zz < S.d && yy >= 0 && yy < S.h && xx >= 0 && xx < S.w)
                    rb[v] = *(const bf16x8*)(x + ((((long)nq * S.d + zz) * S.h + yy) * S.w + xx) * cin + ci);
            }
        }
    };
    auto store = [&](int buf, const bf16x8& ra, const bf16x8 (&rb)[2]) {
        if (tid < 128) {
            const int pl = tid >> 2, c = 8 * (tid & 3);
#pragma unroll
            for (int j = 0; j < 8; ++j) As[buf][c + j][pl] = ra[j];
        }
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int u = tid + NT * v;
            const int pl = u >> 4, kl = 8 * (u & 15);
#pragma unroll
            for (int j = 0; j < 8; ++j) Bs[buf][kl + j][pl] = rb[v][j];
        }
    };
    f32x4 acc[2][2];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int g = 0; g < 2; ++g) acc[h][g] = f32x4{0.f, 0.f, 0.f, 0.f};
    bf16x8 ra, rb[2];
    if (ntile > 0) {
        load(pbeg, ra, rb);
        store(0, ra, rb);
    }
    __syncthreads();
    for (int it = 0; it < ntile; ++it) {
        const int buf = it & 1;
        const bool more = it + 1 < ntile;
        if (more) load(pbeg + 32L * (it + 1), ra, rb);
        bf16x8 fa[2], fb[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) fa[h] = *(const bf16x8*)&As[buf][h * 16 + r][q * 8];
#pragma unroll
        for (int g = 0; g < 2; ++g) fb[g] = *(const bf16x8*)&Bs[buf][wv * 32 + g * 16 + r][q * 8];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int g = 0; g < 2; ++g) acc[h][g] = mfma(fa[h], fb[g], acc[h][g]);
        if (more) store(buf ^ 1, ra, rb);
        __syncthreads();
    }
    const int kk0 = kb0 + wv * 32;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            const int kcol = kk0 + g * 16 + r;
            if (kcol >= K) continue;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int co = co0 + h * 16 + q * 4 + t;
                if (co >= cout) continue;
                if (ORDERED)
                    dw[((long)blockIdx.z * cout + co) * K + kcol] = acc[h][g][t];
                else
                    atomicAdd(&dw[(long)co * K + kcol], acc[h][g][t]);
            }
        }
}

// positions per chunk of the weight- and bias-gradient sums
constexpr long DW_CHUNK = 4096, DB_CHUNK = 8192;

// bias gradient: db[co] += sum over positions of dy[p][co]; ORDERED: the chunk's sum, its eight partial sums
// added in a fixed order, is stored to db[chunk * cout + co]
template <bool ORDERED>
__global__ void __launch_bounds__(NT) k_bias_grad(const __bf16* __restrict__ dy, float* __restrict__ db, long M, int cout,
                                                  long chunk) {
    __shared__ float part[NT];
    const int co = threadIdx.x % 32 + blockIdx.y * 32;
    const int lane_p = threadIdx.x / 32;
    const long pbeg = (long)blockIdx.x * chunk;
    const long pend = pbeg + chunk < M ? pbeg + chunk : M;
    float s = 0.f;
    if (co < cout)
        for (long p = pbeg + lane_p; p < pend; p += NT / 32) s += bf2f(dy[p * cout + co]);
    if (!ORDERED) {
        if (co < cout) atomicAdd(&db[co], s);
        return;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (lane_p == 0 && co < cout) {
        float t = part[threadIdx.x];
        for (int k = 1; k < NT / 32; ++k) t += part[threadIdx.x + 32 * k];
        db[(long)blockIdx.x * cout + co] = t;
    }
}
// out[e] += sum over z = 0 .. nz-1 of part[z * ne + e], in that order
__global__ void __launch_bounds__(NT) k_sum_slices(const float* __restrict__ part, float* __restrict__ out, long ne, int nz) {
    const long e = (long)blockIdx.x * NT + threadIdx.x;
    if (e >= ne) return;
    float s = 0.f;
    for (int z = 0; z < nz; ++z) s += part[(long)z * ne + e];
    out[e] += s;
}

__global__ void k_relu_bwd(__bf16* __restrict__ g, const __bf16* __restrict__ y, long n) {
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i < n && !(bf2f(y[i]) > 0.f)) g[i] = f2bf(0.f);
}

// 2x average pooling (fine shape S, coarse S/2) and its adjoint; nearest 2x upsampling and its adjoint
__global__ void k_pool2(const __bf16* __restrict__ x, __bf16* __restrict__ y, Shape S, int c, int bwd) {
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    const int D = S.d / 2, H = S.h / 2, W = S.w / 2;
    const long n = (long)S.n * D * H * W * c;
    if (i >= n) return;
    const int ch = (int)(i % c);
    long t = i / c;
    const int xx = (int)(t % W);
    t /= W;
    const int yy = (int)(t % H);
    t /= H;
    const int zz = (int)(t % D);
    const int nn = (int)(t / D);
    float s = 0.f;
    for (int o = 0; o < 8; ++o) {
        const long f = ((((long)nn * S.d + 2 * zz + (o >> 2)) * S.h + 2 * yy + ((o >> 1) & 1)) * S.w + 2 * xx + (o & 1)) * c + ch;
        if (!bwd) s += bf2f(x[f]);
        else y[f] = f2bf(0.125f * bf2f(x[i]));   // x: coarse gradient, y: fine gradient
    }
    if (!bwd) y[i] = f2bf(0.125f * s);
}
__global__ void k_up2(const __bf16* __restrict__ x, __bf16* __restrict__ y, Shape S, int c, int bwd) {
    // S: fine shape.  forward: y(fine) = x(coarse parent); backward: y(coarse) = sum of x(fine children)
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    const int D = S.d / 2, H = S.h / 2, W = S.w / 2;
    const long n = (long)S.n * D * H * W * c;
    if (i >= n) return;
    const int ch = (int)(i % c);
    long t = i / c;
    const int xx = (int)(t % W);
    t /= W;
    const int yy = (int)(t % H);
    t /= H;
    const int zz = (int)(t % D);
    const int nn = (int)(t / D);
    float s = 0.f;
    for (int o = 0; o < 8; ++o) {
        const long f = ((((long)nn * S.d + 2 * zz + (o >> 2)) * S.h + 2 * yy + ((o >> 1) & 1)) * S.w + 2 * xx + (o & 1)) * c + ch;
        if (bwd) s += bf2f(x[f]);
        else y[f] = x[i];
    }
    if (bwd) y[i] = f2bf(s);
}
__global__ void k_add(const __bf16* __restrict__ a, const __bf16* __restrict__ b, __bf16* __restrict__ y, long n) {
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i < n) y[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}
// mean squared error over the first `used` of `c` channels; writes the gradient (BF16) and adds the sum of
// squares to *loss (FP32)
__global__ void __launch_bounds__(NT) k_mse(const __bf16* __restrict__ y, const __bf16* __restrict__ t,
                                            __bf16* __restrict__ g, float* __restrict__ loss, long npos, int c, int used,
                                            float scale) {
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    __shared__ float red[NT];
    float s = 0.f;
    if (i < npos * c) {
        const int ch = (int)(i % c);
        const float d = ch < used ? bf2f(y[i]) - bf2f(t[i]) : 0.f;
        g[i] = f2bf(2.f * scale * d);
        s = d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = NT / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}
// Adam on FP32 master weights (bias-corrected, Kingma & Ba 2015)
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
// FP32 master weights w[co][27 cin] -> BF16 forward rows wf[co][kpad] and the flipped, transposed rows of the
// input-gradient convolution wb[ci][kpadb] (wb[ci][off' cout + co] = w[co][(26 - off') cin + ci])
__global__ void k_weights(const float* __restrict__ w, __bf16* __restrict__ wf, __bf16* __restrict__ wb, int cin,
                          int cout, int kpad, int kpadb) {
    const long i = (long)blockIdx.x * NT + threadIdx.x;
    const long nf = (long)cout * kpad, nb = (long)cin * kpadb;
    if (i < nf) {
        const int co = (int)(i / kpad), k = (int)(i % kpad);
        wf[i] = f2bf(k < 27 * cin ? w[(long)co * 27 * cin + k] : 0.f);
    }
    if (wb && i < nb) {
        const int ci = (int)(i / kpadb), k = (int)(i % kpadb);
        float v = 0.f;
        if (k < 27 * cout) {
            const int off = k / cout, co = k % cout;
            v = w[(long)co * 27 * cin + (26 - off) * cin + ci];
        }
        wb[i] = f2bf(v);
    }
}

inline unsigned blocks(long n) { return (unsigned)((n + NT - 1) / NT); }
inline int launched(hipStream_t) {
    return hipGetLastError() == hipSuccess ? MFLBM_OK : MFLBM_ERR_HIP;
}
inline int kpad_of(int cin) { return (27 * cin + 31) / 32 * 32; }

}  // namespace

extern "C" int mflbm_nn_mfma_probe(const void* a, const void* b, float* d, void* stream) {
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, (const __bf16*)a, (const __bf16*)b, d);
    return launched((hipStream_t)stream);
}

extern "C" int mflbm_nn_conv3d_fwd(const void* x, const void* wf, const float* bias, void* y, int32_t n, int32_t d,
                                   int32_t h, int32_t w, int32_t cin, int32_t cout, int32_t relu, void* stream) {
    if (!x || !wf || !y || n < 1 || d < 1 || h < 1 || w < 1 || cin < 8 || cin % 8 || cout < 1) return MFLBM_ERR_ARG;
    const Shape S{n, d, h, w};
    const long M = S.count();
    const dim3 grid((unsigned)((M + 127) / 128), (unsigned)((cout + 31) / 32));
    if (relu)
        hipLaunchKernelGGL(k_conv_fwd<1>, grid, dim3(NT), 0, (hipStream_t)stream, (const __bf16*)x, (const __bf16*)wf,
                           bias, (__bf16*)y, S, (int)cin, (int)cout, kpad_of(cin));
    else
        hipLaunchKernelGGL(k_conv_fwd<0>, grid, dim3(NT), 0, (hipStream_t)stream, (const __bf16*)x, (const __bf16*)wf,
                           bias, (__bf16*)y, S, (int)cin, (int)cout, kpad_of(cin));
    return launched((hipStream_t)stream);
}

extern "C" int mflbm_nn_conv3d_dw(const void* x, const void* dy, float* dw, float* dbias, int32_t n, int32_t d, int32_t h,
                                  int32_t w, int32_t cin, int32_t cout, void* stream) {
    if (!x || !dy || !dw || n < 1 || d < 1 || h < 1 || w < 8 || w % 8 || cin < 8 || cin % 8 || cout < 8 || cout % 8 ||
        (long)n * d * h * w >= (1L << 31))
        return MFLBM_ERR_ARG;
    const Shape S{n, d, h, w};
    const long M = S.count();
    const dim3 grid((unsigned)((cout + 31) / 32), (unsigned)((27 * cin + 127) / 128), (unsigned)((M + DW_CHUNK - 1) / DW_CHUNK));
    hipLaunchKernelGGL(k_conv_dw<false>, grid, dim3(NT), 0, (hipStream_t)stream, (const __bf16*)x, (const __bf16*)dy, dw,
                       S, (int)cin, (int)cout, DW_CHUNK);
    if (dbias)
        hipLaunchKernelGGL(k_bias_grad<false>, dim3((unsigned)((M + DB_CHUNK - 1) / DB_CHUNK), (unsigned)((cout + 31) / 32)),
                           dim3(NT), 0, (hipStream_t)stream, (const __bf16*)dy, dbias, M, (int)cout, DB_CHUNK);
    return launched((hipStream_t)stream);
}

extern "C" int64_t mflbm_nn_conv3d_dw_work(int32_t n, int32_t d, int32_t h, int32_t w, int32_t cin, int32_t cout) {
    if (n < 1 || d < 1 || h < 1 || w < 1 || cin < 1 || cout < 1) return -1;
    const long M = (long)n * d * h * w;
    return ((M + DW_CHUNK - 1) / DW_CHUNK) * cout * 27L * cin + ((M + DB_CHUNK - 1) / DB_CHUNK) * cout;
}

namespace {

// The launches of the fixed-order weight and bias gradients on a workspace of at least
// mflbm_nn_conv3d_dw_work(...) floats.  Every workspace element the sums read is written first by the same
// call (each chunk slice of the weight and bias partials is stored whole), so the workspace needs no clearing.
int dw_ordered_launch(const void* x, const void* dy, float* dw, float* dbias, float* work, const Shape& S, int cin,
                      int cout, hipStream_t st) {
    const long M = S.count();
    const long K = 27L * cin, nzw = (M + DW_CHUNK - 1) / DW_CHUNK, nzb = (M + DB_CHUNK - 1) / DB_CHUNK;
    float* bpart = work + nzw * cout * K;
    const dim3 grid((unsigned)((cout + 31) / 32), (unsigned)((K + 127) / 128), (unsigned)nzw);
    hipLaunchKernelGGL(k_conv_dw<true>, grid, dim3(NT), 0, st, (const __bf16*)x, (const __bf16*)dy, work, S, (int)cin,
                       (int)cout, DW_CHUNK);
    hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((cout * K + NT - 1) / NT)), dim3(NT), 0, st, (const float*)work, dw,
                       cout * K, (int)nzw);
    if (dbias) {
        hipLaunchKernelGGL(k_bias_grad<true>, dim3((unsigned)nzb, (unsigned)((cout + 31) / 32)), dim3(NT), 0, st,
                           (const __bf16*)dy, bpart, M, (int)cout, DB_CHUNK);
        hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((cout + NT - 1) / NT)), dim3(NT), 0, st, (const float*)bpart, dbias,
                           (long)cout, (int)nzb);
    }
    return launched(st);
}

// Workspace of the allocation-free weight gradient: one array per GPU in the library's code object, held until
// the process ends.  Host code never passes its name to a kernel (on the host it is a stand-in, not a device
// address): the device address of each GPU's copy comes from hipGetSymbolAddress once per device, with that
// device current, and only that address is used.
constexpr long DW_SCRATCH_FLOATS = 4L << 20;   // 16 MiB
static __device__ __align__(256) float scratch[DW_SCRATCH_FLOATS];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
float* scratch_ptr[MAX_DEVICES] = {};

// the current device's scratch address, resolved on the first call on that device
int scratch_of_current_device(float** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return MFLBM_ERR_NODEV;
    if (dev < 0 || dev >= MAX_DEVICES) return MFLBM_ERR_ARG;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MFLBM_ERR_HIP;
        scratch_ptr[dev] = (float*)p;
    }
    *out = scratch_ptr[dev];
    return MFLBM_OK;
}

bool dw_args_ok(const void* x, const void* dy, const float* dw, int32_t n, int32_t d, int32_t h, int32_t w, int32_t cin,
                int32_t cout) {
    return x && dy && dw && n >= 1 && d >= 1 && h >= 1 && w >= 8 && w % 8 == 0 && cin >= 8 && cin % 8 == 0 && cout >= 8 &&
           cout % 8 == 0 && (long)n * d * h * w < (1L << 31);
}

}  // namespace

extern "C" int mflbm_nn_conv3d_dw_ordered(const void* x, const void* dy, float* dw, float* dbias, float* work,
                                          int64_t work_floats, int32_t n, int32_t d, int32_t h, int32_t w, int32_t cin,
                                          int32_t cout, void* stream) {
    if (!work || !dw_args_ok(x, dy, dw, n, d, h, w, cin, cout)) return MFLBM_ERR_ARG;
    if (work_floats < mflbm_nn_conv3d_dw_work(n, d, h, w, cin, cout)) return MFLBM_ERR_ARG;
    return dw_ordered_launch(x, dy, dw, dbias, work, Shape{n, d, h, w}, cin, cout, (hipStream_t)stream);
}

extern "C" int64_t mflbm_nn_conv3d_dw_scratch_floats(void) { return DW_SCRATCH_FLOATS; }

extern "C" int mflbm_nn_conv3d_dw_ordered_noalloc(const void* x, const void* dy, float* dw, float* dbias, int32_t n,
                                                  int32_t d, int32_t h, int32_t w, int32_t cin, int32_t cout,
                                                  void* stream) {
    if (!dw_args_ok(x, dy, dw, n, d, h, w, cin, cout)) return MFLBM_ERR_ARG;
    if (mflbm_nn_conv3d_dw_work(n, d, h, w, cin, cout) > DW_SCRATCH_FLOATS) return MFLBM_ERR_ARG;
    float* work = nullptr;
    const int rc = scratch_of_current_device(&work);
    if (rc) return rc;
    return dw_ordered_launch(x, dy, dw, dbias, work, Shape{n, d, h, w}, cin, cout, (hipStream_t)stream);
}

extern "C" int mflbm_nn_weights_to_bf16(const float* w, void* wf, void* wb, int32_t cin, int32_t cout, void* stream) {
    if (!w || !wf || cin < 1 || cout < 1) return MFLBM_ERR_ARG;
    const int kp = kpad_of(cin), kpb = kpad_of(cout);
    const long n = (long)cout * kp > (long)cin * kpb ? (long)cout * kp : (long)cin * kpb;
    hipLaunchKernelGGL(k_weights, dim3(blocks(n)), dim3(NT), 0, (hipStream_t)stream, w, (__bf16*)wf, (__bf16*)wb, (int)cin,
                       (int)cout, kp, kpb);
    return launched((hipStream_t)stream);
}

extern "C" int mflbm_nn_relu_bwd(void* g, const void* y, int64_t n, void* stream) {
    if (!g || !y || n < 0) return MFLBM_ERR_ARG;
    if (n) hipLaunchKernelGGL(k_relu_bwd, dim3(blocks(n)), dim3(NT), 0, (hipStream_t)stream, (__bf16*)g, (const __bf16*)y, (long)n);
    return launched((hipStream_t)stream);
}

// Input features of the warm-start network: the fluid indicator, d / 8 and (d / 8)^2 with d the Euclidean
// distance from a node to the nearest solid node capped at 16, in channels 0..2 of a channels-last BF16 tensor.
// The squared distance is a separable minimum over offsets of at most 16 per axis (exact below the cap: the
// nearest solid within 16 lies inside that cube), wrapped along periodic axes and cut at the others.
constexpr int DCAP = 16;
constexpr int DBIG = 1 << 20;
__device__ __forceinline__ int wrap_or_cut(int a, int n, int per) { return per ? ((a % n) + n) % n : (a >= 0 && a < n ? a : -1); }
__global__ void k_dist_x(const int8_t* __restrict__ fluid, int* __restrict__ g, int d, int h, int w, int per) {
    const long t = (long)blockIdx.x * NT + threadIdx.x;
    if (t >= (long)d * h * w) return;
    const int x = (int)(t % w);
    const long row = t - x;
    int best = DBIG;
    for (int a = -DCAP; a <= DCAP; ++a) {
        const int xx = wrap_or_cut(x + a, w, per);
        if (xx >= 0 && fluid[row + xx] == 0) best = min(best, a * a);
    }
    g[t] = best;
}
__global__ void k_dist_y(const int* __restrict__ gin, int* __restrict__ gout, int d, int h, int w, int per) {
    const long t = (long)blockIdx.x * NT + threadIdx.x;
    if (t >= (long)d * h * w) return;
    const int x = (int)(t % w), y = (int)((t / w) % h);
    const long base = t - (long)y * w;
    int best = DBIG;
    for (int a = -DCAP; a <= DCAP; ++a) {
        const int yy = wrap_or_cut(y + a, h, per);
        if (yy >= 0) best = min(best, gin[base + (long)yy * w] + a * a);
    }
    gout[t] = best;
    (void)x;
}
__global__ void k_dist_z_features(const int8_t* __restrict__ fluid, const int* __restrict__ gin, __bf16* __restrict__ out,
                                  int d, int h, int w, int cin, int per) {
    const long t = (long)blockIdx.x * NT + threadIdx.x;
    const long plane = (long)h * w;
    if (t >= (long)d * plane) return;
    const int z = (int)(t / plane);
    const long base = t - (long)z * plane;
    int best = DBIG;
    for (int a = -DCAP; a <= DCAP; ++a) {
        const int zz = wrap_or_cut(z + a, d, per);
        if (zz >= 0) best = min(best, gin[base + (long)zz * plane] + a * a);
    }
    const double dist = fmin(sqrt((double)best), (double)DCAP) / 8.0;
    __bf16* o = out + t * cin;
    o[0] = f2bf((float)fluid[t]);
    o[1] = f2bf((float)dist);
    o[2] = f2bf((float)(dist * dist));
}
extern "C" int mflbm_nn_distance_features(const int8_t* fluid, int32_t* work, void* x, int32_t d, int32_t h, int32_t w,
                                          int32_t cin, int32_t zper, int32_t yper, int32_t xper, void* stream) {
    if (!fluid || !work || !x || d < 1 || h < 1 || w < 1 || cin < 3) return MFLBM_ERR_ARG;
    const long cnt = (long)d * h * w;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_dist_x, dim3(blocks(cnt)), dim3(NT), 0, st, fluid, (int*)work, (int)d, (int)h, (int)w, (int)xper);
    hipLaunchKernelGGL(k_dist_y, dim3(blocks(cnt)), dim3(NT), 0, st, (const int*)work, (int*)work + cnt, (int)d, (int)h,
                       (int)w, (int)yper);
    hipLaunchKernelGGL(k_dist_z_features, dim3(blocks(cnt)), dim3(NT), 0, st, fluid, (const int*)work + cnt, (__bf16*)x,
                       (int)d, (int)h, (int)w, (int)cin, (int)zper);
    return launched(st);
}

extern "C" int mflbm_nn_pool2(const void* x, void* y, int32_t n, int32_t d, int32_t h, int32_t w, int32_t c, int32_t bwd,
                              void* stream) {
    if (!x || !y || d % 2 || h % 2 || w % 2 || c < 1) return MFLBM_ERR_ARG;
    const Shape S{n, d, h, w};
    const long cnt = (long)n * (d / 2) * (h / 2) * (w / 2) * c;
    hipLaunchKernelGGL(k_pool2, dim3(blocks(cnt)), dim3(NT), 0, (hipStream_t)stream, (const __bf16*)x, (__bf16*)y, S, (int)c,
                       (int)bwd);
    return launched((hipStream_t)stream);
}

extern "C" int mflbm_nn_up2(const void* x, void* y, int32_t n, int32_t d, int32_t h, int32_t w, int32_t c, int32_t bwd,
                            void* stream) {
    if (!x || !y || d % 2 || h % 2 || w % 2 || c < 1) return MFLBM_ERR_ARG;
    const Shape S{n, d, h, w};
    const long cnt = (long)n * (d / 2) * (h / 2) * (w / 2) * c;
    hipLaunchKernelGGL(k_up2, dim3(blocks(cnt)), dim3(NT), 0, (hipStream_t)stream, (const __bf16*)x, (__bf16*)y, S, (int)c,
                       (int)bwd);
    return launched((hipStream_t)stream);
}

extern "C" int mflbm_nn_add(const void* a, const void* b, void* y, int64_t n, void* stream) {
    if (!a || !b || !y || n < 0) return MFLBM_ERR_ARG;
    if (n) hipLaunchKernelGGL(k_add, dim3(blocks(n)), dim3(NT), 0, (hipStream_t)stream, (const __bf16*)a, (const __bf16*)b,
                              (__bf16*)y, (long)n);
    return launched((hipStream_t)stream);
}

extern "C" int mflbm_nn_mse(const void* y, const void* t, void* g, float* loss, int64_t npos, int32_t c, int32_t used,
                            float scale, void* stream) {
    if (!y || !t || !g || !loss || npos < 0 || c < 1 || used < 0 || used > c) return MFLBM_ERR_ARG;
    const long n = (long)npos * c;
    if (n) hipLaunchKernelGGL(k_mse, dim3(blocks(n)), dim3(NT), 0, (hipStream_t)stream, (const __bf16*)y, (const __bf16*)t,
                              (__bf16*)g, loss, (long)npos, (int)c, (int)used, scale);
    return launched((hipStream_t)stream);
}

extern "C" int mflbm_nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                             float eps, int32_t step, void* stream) {
    if (!w || !g || !m || !v || n < 0 || step < 1) return MFLBM_ERR_ARG;
    float c1 = 1.f, c2 = 1.f;
    for (int s = 0; s < step; ++s) {
        c1 *= b1;
        c2 *= b2;
    }
    if (n) hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(NT), 0, (hipStream_t)stream, w, g, m, v, (long)n, lr, b1, b2,
                              eps, 1.f - c1, 1.f - c2);
    return launched((hipStream_t)stream);
}
