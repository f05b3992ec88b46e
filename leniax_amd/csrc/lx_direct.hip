// Direct-convolution potential (leniax/core.py:105-146 get_potential): the depthwise
// cross-correlation of the wrap-padded state with small real kernels, as
// jax.lax.conv_general_dilated(..., 'VALID', feature_group_count=C) computes it, with the
// padding of leniax/helpers.py:465-481 (before = size // 2 on every axis).
//
//   U[b][k][y][x] = sum_{i<kh, j<kw} A[b][cin k][(y + i - ph) mod H][(x + j - pw) mod W] * kern[k][i][j]
//
// One workgroup computes a TH x TW tile of one (world, kernel) output: the input tile and its
// halo are staged in LDS (wrap indexing), each thread keeps a 1 x 4 strip of outputs and slides
// a 4-wide register window along each kernel row, so that one LDS read feeds 4 FMAs; the kernel
// taps are uniform (scalar loads).  Any world size; kernels up to 64 x 64.
#include "lx_internal.h"

namespace {

constexpr int TH = 16, TW = 64, MAXK = 64;

// LDS tile row length: the output tile, the kernel's halo and one spare column (the window's
// last shift reads it), rounded up to a multiple of 4 so that rows start 16-byte aligned
__host__ __device__ inline int tile_cols(int kw) { return (TW + kw + 3) & ~3; }

// Stage rows [y0 - qh, y0 - qh + trows) x columns [x0 - qw, x0 - qw + tcols) of an H x W
// image, wrapped, into tile[trows][tcols]: each thread keeps one column (one modulo per thread,
// not per element) and walks rows nsplit apart.  tcols <= 256.
LX_DEV void stage_tile(float* tile, const float* img, int H, int W, int y0, int x0, int qh, int qw, int trows,
                       int tcols) {
  const int nsplit = 256 / tcols;
  const int t = threadIdx.x;
  if (t < nsplit * tcols) {
    const int c = t % tcols, r0 = t / tcols;
    int gx = (x0 + c - qw) % W;
    gx += gx < 0 ? W : 0;
    int gy = (y0 + r0 - qh) % H;
    gy += gy < 0 ? H : 0;
    // 8 rows per round: the loads are all issued before the first LDS write waits on them.
    // The row advances by nsplit, i.e. by step = nsplit mod H < H, so one conditional
    // subtraction keeps gy in [0, H) for any world height.
    const int step = nsplit % H;
    for (int rb = r0; rb < trows; rb += 8 * nsplit) {
      float v[8];
      int rows[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        rows[u] = rb + u * nsplit;
        v[u] = rows[u] < trows ? img[int64_t(gy) * W + gx] : 0.f;
        gy += step;
        gy -= gy >= H ? H : 0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (rows[u] < trows) tile[rows[u] * tcols + c] = v[u];
    }
  }
}

// acc[e] += sum_{j < kw} k_j row[e + j], e < 4, j in increasing order (k_j = kr[j], or kr[kw-1-j]
// when FLIP): taps in blocks of 4 against an 8-wide window read 16 bytes at a time (16 FMAs per
// two LDS reads, no register shifts), the remaining taps one at a time.
template <bool FLIP>
LX_DEV void row_taps(const float* row, const float* kr, int kw, float (&acc)[4]) {
  float4 lo = *reinterpret_cast<const float4*>(row);
  int j = 0;
  for (; j + 4 <= kw; j += 4) {
    const float4 hi = *reinterpret_cast<const float4*>(row + j + 4);
    const float w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float kv = FLIP ? kr[kw - 1 - (j + t)] : kr[j + t];
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] = fmaf(kv, w[e + t], acc[e]);
    }
    lo = hi;
  }
  float w0 = lo.x, w1 = lo.y, w2 = lo.z, w3 = lo.w;
  for (; j < kw; ++j) {
    const float kv = FLIP ? kr[kw - 1 - j] : kr[j];
    acc[0] = fmaf(kv, w0, acc[0]);
    acc[1] = fmaf(kv, w1, acc[1]);
    acc[2] = fmaf(kv, w2, acc[2]);
    acc[3] = fmaf(kv, w3, acc[3]);
    w0 = w1; w1 = w2; w2 = w3; w3 = row[j + 4];
  }
}

struct DirectArgs {
  const float* A;
  const float* kern;
  float* U;
  int H, W, C, K, kh, kw, ph, pw;
  int cin[LX_MAX_K];
};

__global__ __launch_bounds__(256) void direct_kernel(DirectArgs a) {
  extern __shared__ __align__(16) float tile[];  // (TH + kh - 1) rows x tile_cols(kw) columns
  const int tcols = tile_cols(a.kw);
  const int trows = TH + a.kh - 1;
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
  const int64_t bk = blockIdx.z;
  const int64_t b = bk / a.K;
  const int k = int(bk - b * a.K);
  const float* A = a.A + (b * a.C + a.cin[k]) * int64_t(a.H) * a.W;
  stage_tile(tile, A, a.H, a.W, y0, x0, a.ph, a.pw, trows, tcols);
  __syncthreads();
  const int ty = threadIdx.x / (TW / 4), tx4 = (threadIdx.x % (TW / 4)) * 4;
  const float* kk = a.kern + int64_t(k) * a.kh * a.kw;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int i = 0; i < a.kh; ++i) row_taps<false>(tile + (ty + i) * tcols + tx4, kk + i * a.kw, a.kw, acc);
  const int y = y0 + ty;
  if (y < a.H) {
    float* out = a.U + (bk * a.H + y) * int64_t(a.W) + x0 + tx4;
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (x0 + tx4 + e < a.W) out[e] = acc[e];
  }
}

// Reverse of the direct potential, state side: dA[b][c] = sum over kernels k fed by c of the
// correlation of dU[b][k] with the flipped kernel, wrap-padded with kh - 1 - ph before (grid:
// tiles x B * C).  The tile of each kernel's dU is staged in LDS in turn.
__global__ __launch_bounds__(256) void direct_adj_kernel(DirectArgs a, const float* dU, float* dA) {
  extern __shared__ __align__(16) float tile[];
  const int tcols = tile_cols(a.kw);
  const int trows = TH + a.kh - 1;
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
  const int64_t bc = blockIdx.z;
  const int64_t b = bc / a.C;
  const int c = int(bc - b * a.C);
  const int qh = a.kh - 1 - a.ph, qw = a.kw - 1 - a.pw;
  const int ty = threadIdx.x / (TW / 4), tx4 = (threadIdx.x % (TW / 4)) * 4;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int k = 0; k < a.K; ++k) {
    if (a.cin[k] != c) continue;
    const float* G = dU + (b * a.K + k) * int64_t(a.H) * a.W;
    __syncthreads();  // the previous kernel's tile is consumed
    stage_tile(tile, G, a.H, a.W, y0, x0, qh, qw, trows, tcols);
    __syncthreads();
    const float* kk = a.kern + int64_t(k) * a.kh * a.kw;
    for (int i = 0; i < a.kh; ++i)  // flipped rows and columns
      row_taps<true>(tile + (ty + i) * tcols + tx4, kk + (a.kh - 1 - i) * a.kw, a.kw, acc);
  }
  const int y = y0 + ty;
  if (y < a.H) {
    float* out = dA + (bc * a.H + y) * int64_t(a.W) + x0 + tx4;
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (x0 + tx4 + e < a.W) out[e] = acc[e];
  }
}

// Reverse of the direct potential, kernel side: dK[k][i][j] = sum over worlds and cells of
// dU[b][k][y][x] A[b][cin k][(y + i - ph) mod H][(x + j - pw) mod W] (grid: kh * kw x K; one block
// per tap, so the sum needs no atomics and is deterministic).
__global__ __launch_bounds__(256) void direct_kgrad_kernel(DirectArgs a, int64_t B, const float* dU, float* dK) {
  const int tap = blockIdx.x;
  const int i = tap / a.kw, j = tap - i * a.kw;
  const int k = blockIdx.y;
  const int64_t n = int64_t(a.H) * a.W;
  float acc = 0.f;
  for (int64_t b = 0; b < B; ++b) {
    const float* G = dU + (b * a.K + k) * n;
    const float* A = a.A + (b * a.C + a.cin[k]) * n;
    for (int64_t p = threadIdx.x; p < n; p += 256) {
      const int y = int(p / a.W), x = int(p - int64_t(y) * a.W);
      int gy = (y + i - a.ph) % a.H;
      gy += gy < 0 ? a.H : 0;
      int gx = (x + j - a.pw) % a.W;
      gx += gx < 0 ? a.W : 0;
      acc = fmaf(G[p], A[int64_t(gy) * a.W + gx], acc);
    }
  }
  __shared__ float red[4];
  const float s = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) dK[(int64_t(k) * a.kh + i) * a.kw + j] = red[0] + red[1] + red[2] + red[3];
}

DirectArgs direct_args(const lx_rule* rule, const float* A, const float* kern, int kh, int kw) {
  DirectArgs a;
  a.A = A; a.kern = kern; a.U = nullptr;
  a.H = rule->dims[0]; a.W = rule->dims[1]; a.C = rule->C; a.K = rule->K;
  a.kh = kh; a.kw = kw; a.ph = kh / 2; a.pw = kw / 2;
  for (int k = 0; k < LX_MAX_K; ++k) a.cin[k] = k < rule->K ? rule->cin[k] : 0;
  return a;
}

}  // namespace

extern "C" int lx_potential_direct_adjoint(const lx_rule* rule, const float* state_d, int64_t B,
                                           const float* kern_d, int32_t kh, int32_t kw, const float* dpot_d,
                                           float* dstate_d, float* dkern_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (rule->ndim != 2) return lx_fail(LX_ENOTSUP, "lx_potential_direct_adjoint: 2-D worlds only");
  if (kh < 1 || kw < 1 || kh > MAXK || kw > MAXK) return lx_fail(LX_ENOTSUP, "kernel %dx%d (1..%d)", kh, kw, MAXK);
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  if (!kern_d || !dpot_d || (!dstate_d && !dkern_d) || (dkern_d && !state_d)) return lx_fail(LX_EINVAL, "NULL pointer");
  if (B * rule->C > 65535) return lx_fail(LX_ENOTSUP, "lx_potential_direct_adjoint: B * C > 65535");
  const DirectArgs a = direct_args(rule, state_d, kern_d, kh, kw);
  hipStream_t st = (hipStream_t)stream;
  if (dstate_d) {
    const size_t lds = size_t(TH + kh - 1) * size_t(tile_cols(kw)) * sizeof(float);
    const dim3 grid(unsigned((a.W + TW - 1) / TW), unsigned((a.H + TH - 1) / TH), unsigned(B * rule->C));
    hipLaunchKernelGGL(direct_adj_kernel, grid, dim3(256), lds, st, a, dpot_d, dstate_d);
  }
  if (dkern_d) {
    hipLaunchKernelGGL(direct_kgrad_kernel, dim3(unsigned(kh * kw), unsigned(rule->K)), dim3(256), 0, st, a, B,
                       dpot_d, dkern_d);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LX_OK : lx_fail(LX_EDEVICE, "direct adjoint launch: %s", hipGetErrorString(e));
}

extern "C" int lx_potential_direct(const lx_rule* rule, const float* state_d, int64_t B, const float* kern_d,
                                   int32_t kh, int32_t kw, float* potential_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (rule->ndim != 2) return lx_fail(LX_ENOTSUP, "lx_potential_direct: 2-D worlds only");
  if (kh < 1 || kw < 1 || kh > MAXK || kw > MAXK) return lx_fail(LX_ENOTSUP, "kernel %dx%d (1..%d)", kh, kw, MAXK);
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  if (!state_d || !kern_d || !potential_d) return lx_fail(LX_EINVAL, "NULL pointer");
  if (B * rule->K > 65535) return lx_fail(LX_ENOTSUP, "lx_potential_direct: B * K > 65535");
  DirectArgs a = direct_args(rule, state_d, kern_d, kh, kw);
  a.U = potential_d;
  const size_t lds = size_t(TH + kh - 1) * size_t(tile_cols(kw)) * sizeof(float);
  const dim3 grid(unsigned((a.W + TW - 1) / TW), unsigned((a.H + TH - 1) / TH), unsigned(B * rule->K));
  hipLaunchKernelGGL(direct_kernel, grid, dim3(256), lds, (hipStream_t)stream, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LX_OK : lx_fail(LX_EDEVICE, "direct potential launch: %s", hipGetErrorString(e));
}
