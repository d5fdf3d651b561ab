// gnn.hip -- kernels of the warm-start graph network (no counterpart in the
// reference): BF16 MFMA matrix products with FP32 accumulation, mean
// aggregation over the cell-adjacency graph and its adjoint, SiLU, masked MSE
// and Adam. The C ABI is declared in include/betse_amd.h; the Python model in
// betse_amd/gnn.py composes the forward and backward passes from these calls.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>

#include "betse_amd.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int EB = 256;   // elementwise block

// float -> bfloat16 bits, round to nearest even (inputs are finite)
__device__ __forceinline__ short to_bf16(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C tile (16 rows x 64 columns per block, one 16x16 tile per wavefront) of
// op(A) op(B) over k in [k_lo, k_hi), with v_mfma_f32_16x16x16bf16_1k.
// Lane l supplies A[r0 + l%16][k0 + 4*(l/16) + j] and B[k0 + 4*(l/16) + j][c0 + l%16],
// j = 0..3, and receives C[r0 + 4*(l/16) + q][c0 + l%16], q = 0..3.
__global__ void __launch_bounds__(256)
k_gemm_bf16(const float* __restrict__ A, long long sam, long long sak,
            const float* __restrict__ B, long long sbk, long long sbn,
            float* __restrict__ out, long long ldo, int m, int n, int kd, int kchunk) {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int r0 = blockIdx.x * 16;
    const int c0 = blockIdx.y * 64 + wave * 16;
    const int k_lo = blockIdx.z * kchunk;
    const int k_hi = min(kd, k_lo + kchunk);
    if (c0 >= n) return;
    const int row = r0 + lane % 16, col = c0 + lane % 16, kq = 4 * (lane / 16);
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = k_lo; k0 < k_hi; k0 += 16) {
        v4s a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + kq + j;
            const bool kin = (k < k_hi);
            a[j] = (kin && row < m) ? to_bf16(A[row * sam + k * sak]) : (short)0;
            b[j] = (kin && col < n) ? to_bf16(B[k * sbk + col * sbn]) : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    float* o = out + (long long)blockIdx.z * m * (long long)ldo;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = r0 + kq + q;
        if (r < m && col < n) o[r * ldo + col] = acc[q];
    }
}

// C (+)= sum of the split-K partials
__global__ void __launch_bounds__(EB)
k_sum_splits(const float* ws, int splits, float* C, long long ldc, int m, int n, int beta) {
    const long long e = (long long)blockIdx.x * EB + threadIdx.x;
    if (e >= (long long)m * n) return;
    const int r = (int)(e / n), c = (int)(e % n);
    float s = 0.f;
    for (int q = 0; q < splits; ++q) s += ws[(long long)q * m * n + (long long)r * n + c];
    C[r * ldc + c] = beta ? C[r * ldc + c] + s : s;
}

__device__ __forceinline__ float silu_f(float z) { return z / (1.f + expf(-z)); }

// Z += b; H = silu(Z) (+ Hres)
__global__ void __launch_bounds__(EB)
k_bias_silu(float* Z, const float* b, float* H, const float* Hres, int n, int f) {
    const long long e = (long long)blockIdx.x * EB + threadIdx.x;
    if (e >= (long long)n * f) return;
    const float z = Z[e] + b[e % f];
    Z[e] = z;
    const float s = silu_f(z);
    H[e] = Hres ? Hres[e] + s : s;
}

__global__ void __launch_bounds__(EB)
k_bias(float* Y, const float* b, int n, int f) {
    const long long e = (long long)blockIdx.x * EB + threadIdx.x;
    if (e < (long long)n * f) Y[e] += b[e % f];
}

// dZ = dH * silu'(Z)
__global__ void __launch_bounds__(EB)
k_silu_bwd(const float* dH, const float* Z, float* dZ, long long count) {
    const long long e = (long long)blockIdx.x * EB + threadIdx.x;
    if (e >= count) return;
    const float z = Z[e];
    const float s = 1.f / (1.f + expf(-z));
    dZ[e] = dH[e] * (s * (1.f + z * (1.f - s)));
}

// out[i] = mean_{j in N(i)} H[j]   (deg clamped to 1)
__global__ void __launch_bounds__(EB)
k_mean_agg(const float* H, long long ldh, const int32_t* ptr, const int32_t* idx, float* out, long long ldo,
           int n, int f) {
    const long long e = (long long)blockIdx.x * EB + threadIdx.x;
    if (e >= (long long)n * f) return;
    const int i = (int)(e / f), c = (int)(e % f);
    const int a = ptr[i], b = ptr[i + 1];
    float s = 0.f;
    for (int q = a; q < b; ++q) s += H[(long long)idx[q] * ldh + c];
    out[(long long)i * ldo + c] = s / (float)max(b - a, 1);
}

// adjoint on a symmetric graph: out[j] (+)= sum_{i in N(j)} G[i] / deg_i
__global__ void __launch_bounds__(EB)
k_mean_agg_T(const float* G, long long ldg, const int32_t* ptr, const int32_t* idx, float* out, long long ldo,
             int n, int f, int accumulate) {
    const long long e = (long long)blockIdx.x * EB + threadIdx.x;
    if (e >= (long long)n * f) return;
    const int j = (int)(e / f), c = (int)(e % f);
    float s = 0.f;
    for (int q = ptr[j]; q < ptr[j + 1]; ++q) {
        const int i = idx[q];
        s += G[(long long)i * ldg + c] / (float)max(ptr[i + 1] - ptr[i], 1);
    }
    float* o = out + (long long)j * ldo + c;
    *o = accumulate ? *o + s : s;
}

// column sums, two passes: per-block partials then a final reduction
__global__ void __launch_bounds__(EB)
k_colsum_part(const float* G, long long ldg, int n, int f, int rows_per_block, float* part) {
    const int c = threadIdx.x;
    if (c >= f) return;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(n, r0 + rows_per_block);
    float s = 0.f;
    for (int r = r0; r < r1; ++r) s += G[(long long)r * ldg + c];
    part[(long long)blockIdx.x * f + c] = s;
}

__global__ void __launch_bounds__(EB)
k_colsum_final(const float* part, int parts, int f, float* out) {
    const int c = blockIdx.x * EB + threadIdx.x;
    if (c >= f) return;
    float s = 0.f;
    for (int q = 0; q < parts; ++q) s += part[(long long)q * f + c];
    out[c] = s;
}

// masked MSE: dY = 2 (Y - T) mask / count, per-block partial losses
__global__ void __launch_bounds__(EB)
k_mse(const float* Y, const float* T, const float* mask, long long count, float inv_n, float* dY, float* part) {
    __shared__ float red[EB];
    const long long e = (long long)blockIdx.x * EB + threadIdx.x;
    float l = 0.f;
    if (e < count) {
        const float d = (Y[e] - T[e]) * mask[e];
        dY[e] = 2.f * d * inv_n;
        l = d * d * inv_n;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = EB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(EB)
k_adam(float* p, const float* g, float* m, float* v, long long count, float lr, float b1, float b2, float eps,
       float c1, float c2) {
    const long long e = (long long)blockIdx.x * EB + threadIdx.x;
    if (e >= count) return;
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.f - b1) * gg;
    const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

inline unsigned blocks(long long n) { return (unsigned)((n + EB - 1) / EB); }

}  // namespace

// Scratch of bt_gnn_colsum_noalloc: the per-block partial sums live in this
// file-scope device array instead of a caller workspace. Each GPU the library
// runs on holds its own copy until the process ends; host code reaches it only
// through the device address hipGetSymbolAddress returns for the current
// device (the array's name on the host is not a device address).
static __device__ __align__(256) float scratch[BT_GNN_SCRATCH_FLOATS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
float* g_scratch[MAX_DEVICES] = {};

// the device address of `scratch` on the current device, looked up once per device
int scratch_address(float** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return BT_EHIP;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return BT_EHIP;
        g_scratch[dev] = static_cast<float*>(ptr);
    }
    *out = g_scratch[dev];
    return BT_OK;
}

// rows per partial-sum block of the column sums: 256, doubled until the
// partials (one row of f per block) fit in ws_len floats
int colsum_rows_per_block(int n, int f, long long ws_len) {
    int rpb = 256;
    int parts = (n + rpb - 1) / rpb;
    while ((long long)parts * f > ws_len) {
        rpb *= 2;
        parts = (n + rpb - 1) / rpb;
    }
    return rpb;
}

}  // namespace

extern "C" {

int bt_gnn_gemm(const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk, int64_t sbn,
                float* C, int64_t ldc, int32_t m, int32_t n, int32_t k, int32_t beta,
                float* ws, int32_t ksplit, void* stream) {
    if (!A || !B || !C || m < 1 || n < 1 || k < 1 || ksplit < 1) return BT_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    if (ksplit > 1 && !ws) return BT_EINVAL;
    int kchunk = (k + ksplit - 1) / ksplit;
    kchunk = (kchunk + 15) / 16 * 16;
    const int splits = (k + kchunk - 1) / kchunk;
    dim3 grid((m + 15) / 16, (n + 63) / 64, splits);
    if (splits == 1 && !beta) {
        hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, A, (long long)sam, (long long)sak, B,
                           (long long)sbk, (long long)sbn, C, (long long)ldc, m, n, k, kchunk);
    } else {
        if (!ws) return BT_EINVAL;
        hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, A, (long long)sam, (long long)sak, B,
                           (long long)sbk, (long long)sbn, ws, (long long)n, m, n, k, kchunk);
        hipLaunchKernelGGL(k_sum_splits, dim3(blocks((long long)m * n)), dim3(EB), 0, s,
                           (const float*)ws, splits, C, (long long)ldc, m, n, beta);
    }
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

int bt_gnn_bias_silu(float* Z, const float* b, float* H, const float* Hres, int32_t n, int32_t f, void* stream) {
    if (!Z || !b || !H || n < 1 || f < 1) return BT_EINVAL;
    hipLaunchKernelGGL(k_bias_silu, dim3(blocks((long long)n * f)), dim3(EB), 0, (hipStream_t)stream,
                       Z, b, H, Hres, n, f);
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

int bt_gnn_bias(float* Y, const float* b, int32_t n, int32_t f, void* stream) {
    if (!Y || !b || n < 1 || f < 1) return BT_EINVAL;
    hipLaunchKernelGGL(k_bias, dim3(blocks((long long)n * f)), dim3(EB), 0, (hipStream_t)stream, Y, b, n, f);
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

int bt_gnn_silu_bwd(const float* dH, const float* Z, float* dZ, int64_t count, void* stream) {
    if (!dH || !Z || !dZ || count < 1) return BT_EINVAL;
    hipLaunchKernelGGL(k_silu_bwd, dim3(blocks(count)), dim3(EB), 0, (hipStream_t)stream, dH, Z, dZ,
                       (long long)count);
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

int bt_gnn_mean_agg(const float* H, int64_t ldh, const int32_t* ptr, const int32_t* idx, float* out, int64_t ldo,
                    int32_t n, int32_t f, void* stream) {
    if (!H || !ptr || !idx || !out || n < 1 || f < 1) return BT_EINVAL;
    hipLaunchKernelGGL(k_mean_agg, dim3(blocks((long long)n * f)), dim3(EB), 0, (hipStream_t)stream,
                       H, (long long)ldh, ptr, idx, out, (long long)ldo, n, f);
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

int bt_gnn_mean_agg_T(const float* G, int64_t ldg, const int32_t* ptr, const int32_t* idx, float* out,
                      int64_t ldo, int32_t n, int32_t f, int32_t accumulate, void* stream) {
    if (!G || !ptr || !idx || !out || n < 1 || f < 1) return BT_EINVAL;
    hipLaunchKernelGGL(k_mean_agg_T, dim3(blocks((long long)n * f)), dim3(EB), 0, (hipStream_t)stream,
                       G, (long long)ldg, ptr, idx, out, (long long)ldo, n, f, accumulate);
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

int bt_gnn_colsum(const float* G, int64_t ldg, int32_t n, int32_t f, float* out, float* ws, int32_t ws_len,
                  void* stream) {
    if (!G || !out || !ws || n < 1 || f < 1 || f > EB) return BT_EINVAL;
    const int rpb = colsum_rows_per_block(n, f, ws_len);
    const int parts = (n + rpb - 1) / rpb;
    hipLaunchKernelGGL(k_colsum_part, dim3(parts), dim3(EB), 0, (hipStream_t)stream, G, (long long)ldg, n, f,
                       rpb, ws);
    hipLaunchKernelGGL(k_colsum_final, dim3((f + EB - 1) / EB), dim3(EB), 0, (hipStream_t)stream,
                       (const float*)ws, parts, f, out);
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

int bt_gnn_colsum_noalloc(const float* G, int64_t ldg, int32_t n, int32_t f, float* out, void* stream) {
    if (!G || !out || n < 1 || f < 1) return BT_EINVAL;
    if (f > EB) return BT_EINVAL;          // at most 256 columns (one block of the final sum)
    float* ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != BT_OK) return rc;
    // the same launches as bt_gnn_colsum with ws = scratch: every partial is
    // written before it is read, so nothing needs clearing between calls
    const int rpb = colsum_rows_per_block(n, f, BT_GNN_SCRATCH_FLOATS);
    const int parts = (n + rpb - 1) / rpb;
    hipLaunchKernelGGL(k_colsum_part, dim3(parts), dim3(EB), 0, (hipStream_t)stream, G, (long long)ldg, n, f,
                       rpb, ws);
    hipLaunchKernelGGL(k_colsum_final, dim3((f + EB - 1) / EB), dim3(EB), 0, (hipStream_t)stream,
                       (const float*)ws, parts, f, out);
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

int bt_gnn_mse(const float* Y, const float* T, const float* mask, int64_t count, float inv_n, float* dY,
               float* part, int32_t part_len, void* stream) {
    if (!Y || !T || !mask || !dY || !part || count < 1) return BT_EINVAL;
    if ((long long)blocks(count) > part_len) return BT_EINVAL;
    hipLaunchKernelGGL(k_mse, dim3(blocks(count)), dim3(EB), 0, (hipStream_t)stream, Y, T, mask,
                       (long long)count, inv_n, dY, part);
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

int bt_gnn_adam(float* p, const float* g, float* m, float* v, int64_t count, float lr, float b1, float b2,
                float eps, int32_t t, void* stream) {
    if (!p || !g || !m || !v || count < 1 || t < 1) return BT_EINVAL;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(count)), dim3(EB), 0, (hipStream_t)stream, p, g, m, v,
                       (long long)count, lr, b1, b2, eps, c1, c2);
    return hipGetLastError() == hipSuccess ? BT_OK : BT_EHIP;
}

}  // extern "C"
