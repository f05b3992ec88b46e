// C ABI entry points of libleniax_amd: argument checks, spectrum preparation,
// dispatch of the fused run.  The plain step lives in lx_plain.hip.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "lx_internal.h"

static thread_local char g_err[512] = "";

int lx_fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char* lx_last_error(void) { return g_err; }
extern "C" int lx_version(void) { return 1; }



int lx_check_rule(const lx_rule* rule) {
  if (!rule) return lx_fail(LX_EINVAL, "rule is NULL");
  if (rule->ndim != 2 && rule->ndim != 3) return lx_fail(LX_EINVAL, "ndim must be 2 or 3 (got %d)", rule->ndim);
  for (int i = 0; i < rule->ndim; ++i)
    if (rule->dims[i] < 1) return lx_fail(LX_EINVAL, "dims[%d] = %d", i, rule->dims[i]);
  if (rule->C < 1 || rule->C > LX_MAX_C) return lx_fail(LX_EINVAL, "C must be in [1, %d] (got %d)", LX_MAX_C, rule->C);
  if (rule->K < 1 || rule->K > LX_MAX_K) return lx_fail(LX_EINVAL, "K must be in [1, %d] (got %d)", LX_MAX_K, rule->K);
  for (int k = 0; k < rule->K; ++k) {
    if (rule->cin[k] < 0 || rule->cin[k] >= rule->C)
      return lx_fail(LX_EINVAL, "cin[%d] = %d out of range", k, rule->cin[k]);
    if (rule->gf[k] < 0 || rule->gf[k] > LX_IDENTITY) return lx_fail(LX_EINVAL, "gf[%d] = %d unknown", k, rule->gf[k]);
  }
  if (rule->state_fn < 0 || rule->state_fn > LX_STATE_SIMPLE)
    return lx_fail(LX_EINVAL, "state_fn %d unknown", rule->state_fn);
  if (rule->k_quadrant != 0 && rule->k_quadrant != 1)
    return lx_fail(LX_EINVAL, "k_quadrant must be 0 or 1 (got %d)", rule->k_quadrant);
  if (rule->k_quadrant && (rule->ndim != 2 || rule->dims[0] % 2))
    return lx_fail(LX_EINVAL, "quadrant spectra need a 2-D rule with an even first dimension");
  return LX_OK;
}

// ------------------------------------------------------------------ spectra
struct SlotList {
  int32_t s[LX_MAX_K];
};

__global__ void prepare_spectra_kernel(const float2* __restrict__ kf, float2* __restrict__ kh, int64_t S, int K,
                                       int maxk_C, int d0, int d1, int d2, SlotList slots) {
  // d0, d1 leading dims (d0 == 1 for 2-D), d2 the last dim; output last dim d2/2+1
  const int wh = d2 / 2 + 1;
  const int64_t per_k = int64_t(d0) * d1 * wh;
  const int64_t total = S * K * per_k;
  const int64_t full_per_slot = int64_t(d0) * d1 * d2;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    int64_t rem = i;
    const int c = int(rem % wh); rem /= wh;
    const int h = int(rem % d1); rem /= d1;
    const int d = int(rem % d0); rem /= d0;
    const int k = int(rem % K);
    const int64_t s = rem / K;
    const float2* base = kf + (s * maxk_C + slots.s[k]) * full_per_slot;
    const float2 v = base[(int64_t(d) * d1 + h) * d2 + c];
    const int dn = (d0 - d) % d0, hn = (d1 - h) % d1, cn = (d2 - c) % d2;
    const float2 u = base[(int64_t(dn) * d1 + hn) * d2 + cn];
    kh[i] = make_float2(0.5f * (v.x + u.x), 0.5f * (v.y - u.y));
  }
}

extern "C" int lx_prepare_spectra(const lx_rule* rule, const void* K_full_d, int64_t S, int32_t maxk,
                                  const int32_t* slot, void* K_half_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!K_full_d || !K_half_d || !slot) return lx_fail(LX_EINVAL, "NULL pointer");
  if (S < 1 || maxk < 1) return lx_fail(LX_EINVAL, "S and maxk must be positive");
  SlotList sl;
  for (int k = 0; k < rule->K; ++k) {
    if (slot[k] < 0 || slot[k] >= rule->C * maxk) return lx_fail(LX_EINVAL, "slot[%d] = %d out of range", k, slot[k]);
    sl.s[k] = slot[k];
  }
  const int d0 = rule->ndim == 3 ? rule->dims[0] : 1;
  const int d1 = rule->ndim == 3 ? rule->dims[1] : rule->dims[0];
  const int d2 = rule->ndim == 3 ? rule->dims[2] : rule->dims[1];
  const int64_t total = S * rule->K * int64_t(d0) * d1 * (d2 / 2 + 1);
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(prepare_spectra_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const float2*>(K_full_d), static_cast<float2*>(K_half_d), S, rule->K,
                     rule->C * maxk, d0, d1, d2, sl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return lx_fail(LX_EDEVICE, "prepare_spectra launch: %s", hipGetErrorString(e));
  return LX_OK;
}

// Real quadrants of half spectra (K_quad[i][u][v] = Re K[i][u][v], u <= H / 2) and how far the
// spectra are from real and even along the first axis: dev[0] = max over (i, u, v) of
// |Im K[u][v]| + |Im K[H - u][v]| + |Re K[u][v] - Re K[H - u][v]| (NaN counts as infinite),
// dev[1] = max |Re K|.  Non-negative floats order as their bit patterns: one unsigned max per wave.
__global__ void quadrant_clear_kernel(unsigned* dev) {
  if (threadIdx.x < 2) dev[threadIdx.x] = 0u;
}

__global__ void quadrant_kernel(const float2* __restrict__ kh, float* __restrict__ kq, int64_t n, int H, int WH,
                                unsigned* dev) {
  const int HQ = H / 2 + 1;
  const int64_t total = n * HQ * int64_t(WH);
  float d = 0.f, mx = 0.f;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int v = int(i % WH);
    const int u = int((i / WH) % HQ);
    const int64_t img = i / (int64_t(HQ) * WH);
    const float2* kb = kh + img * H * WH;
    const float2 a = kb[int64_t(u) * WH + v];
    const float2 b = kb[int64_t((H - u) % H) * WH + v];
    kq[i] = a.x;
    const float e = fabsf(a.y) + fabsf(b.y) + fabsf(a.x - b.x);
    d = (e == e) ? fmaxf(d, e) : __builtin_inff();
    const float m = fabsf(a.x);
    mx = (m == m) ? fmaxf(mx, m) : __builtin_inff();
  }
  for (int o = 32; o > 0; o >>= 1) {
    d = fmaxf(d, __shfl_xor(d, o));
    mx = fmaxf(mx, __shfl_xor(mx, o));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(dev, __float_as_uint(d));
    atomicMax(dev + 1, __float_as_uint(mx));
  }
}

extern "C" int lx_quadrant_spectra(const lx_rule* rule, const void* K_half_d, int64_t n, void* K_quad_d, float* dev_d,
                                   void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (rule->ndim != 2 || rule->dims[0] % 2)
    return lx_fail(LX_EINVAL, "lx_quadrant_spectra: a 2-D rule with an even first dimension");
  if (n < 0) return lx_fail(LX_EINVAL, "n < 0");
  if (n == 0) return LX_OK;
  if (!dev_d || !K_half_d || !K_quad_d) return lx_fail(LX_EINVAL, "bad arguments: NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  unsigned* dev = reinterpret_cast<unsigned*>(dev_d);
  hipLaunchKernelGGL(quadrant_clear_kernel, dim3(1), dim3(64), 0, st, dev);
  const int H = rule->dims[0], WH = rule->dims[1] / 2 + 1;
  const int64_t total = n * (H / 2 + 1) * int64_t(WH);
  const unsigned blocks = unsigned(std::min<int64_t>((total + 255) / 256, 8192));
  hipLaunchKernelGGL(quadrant_kernel, dim3(blocks), dim3(256), 0, st, static_cast<const float2*>(K_half_d),
                     static_cast<float*>(K_quad_d), n, H, WH, dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return lx_fail(LX_EDEVICE, "quadrant_spectra launch: %s", hipGetErrorString(e));
  return LX_OK;
}

// ------------------------------------------------------------------ batched run
namespace {
// the world size the step passes take on every axis (2^k in [16, 512], 3 * 2^k in [48, 384])
bool passes_size_ok(const lx_rule* rule) {
  for (int i = 0; i < rule->ndim; ++i) {
    int n = rule->dims[i];
    if (n % 3 == 0) n /= 3;
    if (n < 16 || n > 512 || (n & (n - 1)) || (rule->dims[i] % 3 == 0 && (n < 16 || n > 128))) return false;
  }
  return true;
}
}  // namespace

extern "C" int lx_run_supported(const lx_rule* rule) {
  if (lx_check_rule(rule)) return 0;
  if (lx_fused2d_supported(rule)) return LX_RUN_FUSED;
  return passes_size_ok(rule) ? LX_RUN_PASSES : 0;
}

extern "C" size_t lx_run_scratch_bytes(int64_t B) { return lx_run_scratch_layout(B < 0 ? 0 : B).total; }

extern "C" int lx_run_scratch_size(const lx_rule* rule, int64_t B, int32_t slots, size_t* bytes) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!bytes || B < 0 || slots < 0) return lx_fail(LX_EINVAL, "bad arguments");
  if (lx_fused2d_supported(rule)) {
    *bytes = lx_run_scratch_layout(B).total;
    return LX_OK;
  }
  return lx_passes_scratch(rule, B, slots, bytes);
}

extern "C" int lx_run(const lx_rule* rule, const lx_run_args* a, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!a) return lx_fail(LX_EINVAL, "args is NULL");
  if (a->B < 0 || a->max_steps < 0 || a->slots < 0) return lx_fail(LX_EINVAL, "B, max_steps and slots must be >= 0");
  if (a->stop_policy < LX_STOP_NONE || a->stop_policy > LX_STOP_MASS)
    return lx_fail(LX_EINVAL, "unknown stop policy %d", a->stop_policy);
  if (!(a->R > 0.f) || !(a->T > 0.f)) return lx_fail(LX_EINVAL, "R and T must be positive");
  const int path = lx_run_supported(rule);
  if (!path)
    return lx_fail(LX_ENOTSUP, "lx_run: world size %dx%dx%d unsupported (2^k in [16, 512], 3*2^k in [48, 384])",
                   rule->dims[0], rule->dims[1], rule->ndim == 3 ? rule->dims[2] : 1);
  if (a->B == 0) return LX_OK;  // no world: pointers may be NULL
  if (!a->cells0_d || !a->cells_d || !a->K_half_d || !a->gf_d || !a->w_d || !a->dt_d || !a->scratch_d)
    return lx_fail(LX_EINVAL, "NULL device pointer in lx_run_args");
  if (path == LX_RUN_FUSED) return lx_fused2d_launch(rule, a, (hipStream_t)stream);
  return lx_passes_run(rule, a, (hipStream_t)stream);
}
