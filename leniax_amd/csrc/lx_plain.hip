// Elementwise and per-world kernels of the step: the field and state functions on their own,
// the per-world statistics with their carry, the heuristic scan, and the pointwise part of the
// step's adjoint.  The FFT passes live in lx_fft.hip.
#include <algorithm>
#include <type_traits>

#include "lx_internal.h"

namespace {

struct RuleDev {
  int C, K;
  int cin[LX_MAX_K];
  int gf[LX_MAX_K];
  int state_fn, average;
};

RuleDev to_dev(const lx_rule* r) {
  RuleDev d;
  d.C = r->C;
  d.K = r->K;
  for (int k = 0; k < LX_MAX_K; ++k) {
    d.cin[k] = k < r->K ? r->cin[k] : 0;
    d.gf[k] = k < r->K ? r->gf[k] : 0;
  }
  d.state_fn = r->state_fn;
  d.average = r->average;
  return d;
}

int64_t cells_of(const lx_rule* r) {
  int64_t n = 1;
  for (int i = 0; i < r->ndim; ++i) n *= r->dims[i];
  return n;
}

// leniax/core.py:163-242: G[c] = sum_k w[c][k] g_k(U_k) (/ sum_k w[c][k])
__global__ void field_kernel(RuleDev rd, const float* __restrict__ pot, int64_t B, int64_t n,
                             const float* __restrict__ gf, int64_t gf_bs, const float* __restrict__ w, int64_t w_bs,
                             float* __restrict__ field) {
  const int64_t total = B * n;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t b = i / n, cell = i % n;
    const float* pg = gf + b * gf_bs;
    const float* pw = w + b * w_bs;
    float g[LX_MAX_K];
    for (int k = 0; k < rd.K; ++k) {
      const float s = pg[2 * k + 1];
      g[k] = growth_rcp(rd.gf[k], pg[2 * k], s, 1.f / s, 1.f / (9.f * (s * s)), pot[(b * rd.K + k) * n + cell]);
    }
    for (int c = 0; c < rd.C; ++c) {
      float acc = 0.f, den = 0.f;
      for (int k = 0; k < rd.K; ++k) {
        acc = add_nc(acc, mul_nc(pw[c * rd.K + k], g[k]));
        den = add_nc(den, pw[c * rd.K + k]);
      }
      field[(b * rd.C + c) * n + cell] = rd.average ? (acc / den) : acc;
    }
  }
}

__global__ void state_kernel(int fn, const float* __restrict__ A, const float* __restrict__ G, int64_t B, int64_t nc,
                             const float* __restrict__ dt, int64_t dt_bs, float* __restrict__ out) {
  const int64_t total = B * nc;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const float d = dt[(i / nc) * dt_bs];
    const float a = A[i], g = G[i];
    float v;
    if (fn == LX_STATE_V2) {
      v = add_nc(mul_nc(a, sub_nc(1.f, d)), mul_nc(d, g));
    } else {
      v = add_nc(a, mul_nc(d, g));
      if (fn == LX_STATE_V1) v = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
    }
    out[i] = v;
  }
}

// leniax/statistics.py:134-205: one thread per world walks the T steps.
__global__ void heuristics_kernel(const float* __restrict__ stats, int T, int64_t B, int C, float* __restrict__ cont,
                                  float* __restrict__ nout) {
  const int64_t b = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (b >= B) return;
  const int ns = LX_NSTAT_BASE + C;
  lx_carry c;
  c.t = 0;
  float cm[LX_MAX_C];
  for (int t = 0; t < T; ++t) {
    const float* s = stats + (int64_t(t) * B + b) * ns;
    for (int ch = 0; ch < C; ++ch) cm[ch] = s[LX_ST_CHANNEL_MASS0 + ch];
    lx_heuristic_step(c, s[LX_ST_MASS], cm, s[LX_ST_MASS_VOLUME], C);
    if (cont) cont[int64_t(t) * B + b] = (float)c.cont;
  }
  if (nout) nout[b] = (float)c.n;
}

__global__ void carry_init_kernel(lx_stat_carry* c, int64_t B) {
  const int64_t b = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (b >= B) return;
  lx_carry z;
  for (int i = 0; i < 3; ++i) { z.shift[i] = 0; z.centroid[i] = 0.f; }
  z.angle = 0.f;
  for (int i = 0; i < LX_MAX_C; ++i) z.init_cm[i] = 0.f;
  z.prev_mass = z.prev_sign = 0.f;
  z.mono = z.vol = z.n = z.t = z.pad = 0;
  z.cont = 1;
  c[b].c = z;
}

unsigned grid_for(int64_t total, int block) {
  int64_t g = (total + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

int launched(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return lx_fail(LX_EDEVICE, "%s launch: %s", what, hipGetErrorString(e));
  return LX_OK;
}

}  // namespace

extern "C" int lx_field(const lx_rule* rule, const float* potential_d, int64_t B, const float* gf_d,
                        int64_t gf_bstride, const float* w_d, int64_t w_bstride, float* field_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  if (!potential_d || !gf_d || !w_d || !field_d) return lx_fail(LX_EINVAL, "NULL pointer");
  const int64_t n = cells_of(rule);
  hipLaunchKernelGGL(field_kernel, dim3(grid_for(B * n, 256)), dim3(256), 0, (hipStream_t)stream, to_dev(rule),
                     potential_d, B, n, gf_d, gf_bstride, w_d, w_bstride, field_d);
  return launched("field");
}

extern "C" int lx_state(const lx_rule* rule, const float* state_d, const float* field_d, int64_t B, const float* dt_d,
                        int64_t dt_bstride, float* out_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  if (!state_d || !field_d || !dt_d || !out_d) return lx_fail(LX_EINVAL, "NULL pointer");
  const int64_t nc = cells_of(rule) * rule->C;
  hipLaunchKernelGGL(state_kernel, dim3(grid_for(B * nc, 256)), dim3(256), 0, (hipStream_t)stream, rule->state_fn,
                     state_d, field_d, B, nc, dt_d, dt_bstride, out_d);
  return launched("state");
}

extern "C" int lx_check_heuristics(const float* stats_d, int32_t T, int64_t B, int32_t C, float* cont_d, float* n_d,
                                   void* stream) {
  if (T < 1 || B < 0 || C < 1 || C > LX_MAX_C) return lx_fail(LX_EINVAL, "bad T/B/C");
  if (B == 0) return LX_OK;
  if (!stats_d) return lx_fail(LX_EINVAL, "NULL stats");  // cont_d and n_d are optional
  hipLaunchKernelGGL(heuristics_kernel, dim3(grid_for(B, 128)), dim3(128), 0, (hipStream_t)stream, stats_d, T, B, C,
                     cont_d, n_d);
  return launched("check_heuristics");
}

extern "C" size_t lx_carry_bytes(void) { return sizeof(lx_stat_carry); }

extern "C" int lx_carry_init(int64_t B, void* carry_d, void* stream) {
  if (B < 0) return lx_fail(LX_EINVAL, "bad carry arguments");
  if (B == 0) return LX_OK;
  if (!carry_d) return lx_fail(LX_EINVAL, "NULL pointer");
  hipLaunchKernelGGL(carry_init_kernel, dim3(grid_for(B, 128)), dim3(128), 0, (hipStream_t)stream,
                     static_cast<lx_stat_carry*>(carry_d), B);
  return launched("carry_init");
}


// ====================================================================== statistics, adjoint
namespace {

// ---------------------------------------------------------------- statistics (one block per world)
struct StatArgs {
  const float* cells;
  const float* field;
  const float* pot;
  lx_stat_carry* carry;
  float* stats;
  int64_t B;
  int ndim;
  int dims[3];
  int C, K;
  float R2, R, dt;
};

// Partial sums of one world's statistics over 1 / LX_STAT_PARTS of its cells (grid: parts x B).
// Coordinates are leniax's re-centred ones (leniax/statistics.py:28-33,64 via
// leniax/utils.py:269-293): (index - shift) mod size - size // 2, tracked incrementally.
__global__ __launch_bounds__(256) void stats_part_kernel(StatArgs a) {
  const int64_t b = blockIdx.y;
  const int part = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  lx_stat_carry* sc = a.carry + b;
  const int d0 = a.ndim == 3 ? a.dims[0] : 1, d1 = a.dims[a.ndim - 2], d2 = a.dims[a.ndim - 1];
  const int n = d0 * d1 * d2;
  // shifts reduced into [0, size): ndim == 2 axes (d1, d2) are carry axes (0, 1), ndim == 3 (0, 1, 2)
  auto modp = [](int v, int m) { v %= m; return v < 0 ? v + m : v; };
  const int s0 = a.ndim == 3 ? modp(sc->c.shift[0], d0) : 0;
  const int s1 = modp(sc->c.shift[a.ndim - 2], d1), s2 = modp(sc->c.shift[a.ndim - 1], d2);
  const int i0 = int(int64_t(n) * part / LX_STAT_PARTS), i1 = int(int64_t(n) * (part + 1) / LX_STAT_PARTS);
  float v[LX_NSUM];
#pragma unroll
  for (int i = 0; i < LX_NSUM; ++i) v[i] = 0.f;
  // v: 0 sA, 1 sG, 2 nA, 3 nG, 4 nU, 5..7 sAy, 8..10 sGy, 11..13 sAyy, 14.. channel masses
  const int adv_z = 256 % d2, adv_r = 256 / d2;
  for (int ch = 0; ch < a.C; ++ch) {
    const float* A = a.cells + (b * a.C + ch) * int64_t(n);
    const float* G = a.field + (b * a.C + ch) * int64_t(n);
    float sc_mass = 0.f;
    const int first = i0 + tid;
    int z = first % d2, row = first / d2;
    int y = row % d1, xq = row / d1;
    for (int i = first; i < i1; i += 256) {
      int zc = z - s2; zc += zc < 0 ? d2 : 0;
      int yc = y - s1; yc += yc < 0 ? d1 : 0;
      int xc = xq - s0; xc += xc < 0 ? d0 : 0;
      const float cz = float(zc - d2 / 2), cy = float(yc - d1 / 2), cx = float(xc - d0 / 2);
      const float av = A[i];
      const float g = G[i];
      const float gp = (g != g) ? g : fmaxf(g, 0.f);
      float co[3];
      if (a.ndim == 3) { co[0] = cx; co[1] = cy; co[2] = cz; } else { co[0] = cy; co[1] = cz; co[2] = 0.f; }
      sc_mass += av;
      v[1] += gp;
      v[2] += av > LX_EPS ? 1.f : 0.f;
      v[3] += gp > LX_EPS ? 1.f : 0.f;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        v[5 + d] = fmaf(av, co[d], v[5 + d]);
        v[8 + d] = fmaf(gp, co[d], v[8 + d]);
        v[11 + d] = fmaf(av * co[d], co[d], v[11 + d]);
      }
      // advance by 256 cells
      z += adv_z;
      int carry_r = adv_r;
      if (z >= d2) { z -= d2; ++carry_r; }
      y += carry_r;
      while (y >= d1) { y -= d1; ++xq; }
    }
    v[0] += sc_mass;
#pragma unroll
    for (int c2 = 0; c2 < LX_MAX_C; ++c2) v[14 + c2] += c2 == ch ? sc_mass : 0.f;
  }
  for (int k = 0; k < a.K; ++k) {
    const float* U = a.pot + (b * a.K + k) * int64_t(n);
    for (int i = i0 + tid; i < i1; i += 256) v[4] += U[i] > LX_EPS ? 1.f : 0.f;
  }
  __shared__ float red[4][LX_NSUM];
  wave_sum_n(v);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < LX_NSUM; ++i) red[wave][i] = v[i];
  }
  __syncthreads();
  if (tid < LX_NSUM) sc->part[part][tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
}

// One thread per world: sums the partial slots in order, finishes the step's statistics and
// carry (leniax/statistics.py:64-126).
__global__ void stats_finish_kernel(StatArgs a) {
  const int64_t b = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (b >= a.B) return;
  lx_stat_carry* sc = a.carry + b;
  float out[LX_NSUM];
#pragma unroll
  for (int i = 0; i < LX_NSUM; ++i) {
    float acc = 0.f;
    for (int p = 0; p < LX_STAT_PARTS; ++p) acc += sc->part[p][i];
    out[i] = acc;
  }
  lx_sums s;
  s.sA = out[0]; s.sG = out[1]; s.nA = out[2]; s.nG = out[3]; s.nU = out[4];
  for (int d = 0; d < 3; ++d) { s.sAy[d] = out[5 + d]; s.sGy[d] = out[8 + d]; s.sAyy[d] = out[11 + d]; }
  for (int ch = 0; ch < LX_MAX_C; ++ch) s.sAc[ch] = ch < a.C ? out[14 + ch] : 0.f;
  int32_t dd[3];
  for (int i = 0; i < 3; ++i) dd[i] = a.dims[i];
  const int64_t n = int64_t(a.ndim == 3 ? a.dims[0] : 1) * a.dims[a.ndim - 2] * a.dims[a.ndim - 1];
  lx_carry c = sc->c;
  lx_stats_finish(s, c, a.ndim, dd, a.C, a.R2, a.R, a.dt, LX_STOP_NONE, 0, float(n * a.C),
                  a.stats + b * (LX_NSTAT_BASE + a.C));
  sc->c = c;
}

// The step's statistics from the row-pair records of lx_step_stats' inverse row pass: one
// kPairThreads block per world sums the records in a fixed order (lx_block_pair_sums) and thread 0
// finishes the step (leniax/statistics.py:64-126) as stats_finish_kernel does.
__global__ __launch_bounds__(kPairThreads) void pair_stats_kernel(StatArgs a, const float* __restrict__ part,
                                                                  int64_t pairs) {
  const int64_t b = blockIdx.x;
  const lx_sums s = lx_block_pair_sums(part + b * int64_t(a.C) * pairs * LX_PAIR_REC, a.C, pairs);
  if (threadIdx.x != 0) return;
  lx_stat_carry* sc = a.carry + b;
  int32_t dd[3];
  for (int i = 0; i < 3; ++i) dd[i] = a.dims[i];
  const int64_t n = int64_t(a.ndim == 3 ? a.dims[0] : 1) * a.dims[a.ndim - 2] * a.dims[a.ndim - 1];
  lx_carry c = sc->c;
  lx_stats_finish(s, c, a.ndim, dd, a.C, a.R2, a.R, a.dt, LX_STOP_NONE, 0, float(n * a.C),
                  a.stats + b * (LX_NSTAT_BASE + a.C));
  sc->c = c;
}

// ---------------------------------------------------------------- adjoint of the step, pointwise part
// Forward (leniax/core.py:163-319):  G_c = sum_k w_ck g_k(U_k) (/ D_c = sum_k w_ck),
//   Z_c = A_c + dt G_c (v1, simple) or A_c (1 - dt) + dt G_c (v2),  A'_c = clip(Z_c, 0, 1) (v1).
// Reverse, with lambda_c = dL/dA'_c and dZ_c = lambda_c * mask_c (mask = 1 for the straight-through
// estimator of leniax/core.py:265-270, v2 and simple; 1[0 < Z < 1] for the true clip):
//   dL/dA_c (direct) = dZ_c (v1, simple) or dZ_c (1 - dt) (v2);  dG_c = dt dZ_c
//   dL/ddt = sum dZ_c G_c (v1, simple) or sum dZ_c (G_c - A_c) (v2)
//   dg_k = sum_c dG_c w_ck / D_c (mean) or sum_c dG_c w_ck (sum)
//   dL/dw_ck = sum dG_c (g_k - G_c) / D_c (mean) or sum dG_c g_k (sum)
//   dL/dU_k = dg_k g_k'(U_k);  dL/dm_k, dL/ds_k = sum dg_k dg/dm, dg/ds
struct AdjArgs {
  const float* A;
  const float* G;
  const float* U;
  const float* lam;
  const float* dfield;  // optional dL/dG added to dG (field output used by the loss)
  const float* dpot_in; // optional dL/dU added to dU (potential output used by the loss)
  const float* gf;
  int64_t gf_bs;
  const float* w;
  int64_t w_bs;
  const float* dt;
  int64_t dt_bs;
  float* lam_out;
  float* dU;
  float* g_gf;
  float* g_w;
  float* g_dt;
  float* part_g;  // per-block partial sums of the growth / weight (and one-kernel) gradients
  float* part_s;  // per-block partial sums of the dt gradient (state kernel)
  float* ysum_g;
  float* ysum_s;
  int64_t cells;
  int C, K;
  int cin[LX_MAX_K];
  int gfk[LX_MAX_K];
  int state_fn, average, ste;
};

LX_DEV float adj_dz(const AdjArgs& a, float A, float G, float lam, float dtw) {
  if (a.state_fn != LX_STATE_V1 || a.ste) return lam;
  const float z = add_nc(A, mul_nc(dtw, G));
  return (z > 0.f && z < 1.f) ? lam : 0.f;
}

template <int NV>
LX_DEV void block_partial(float (&v)[NV], float* part) {
  lx_block_partial<NV>(v, part);
}

// grid (blocks per image, B * K): dL/dU_k and the growth / weight gradients; reads lambda before the
// state kernel overwrites it, so lam_out may alias lam
__global__ __launch_bounds__(256) void adj_growth_kernel(AdjArgs a) {
  const int64_t bk = blockIdx.y;
  const int64_t b = bk / a.K;
  const int k = int(bk % a.K);
  const float* pw = a.w + b * a.w_bs;
  const float* pg = a.gf + b * a.gf_bs;
  const float dtw = a.dt[b * a.dt_bs];
  const float gm = pg[2 * k], gs = pg[2 * k + 1];
  const float inv_s = 1.f / gs, inv_9s2 = 1.f / (9.f * (gs * gs));
  const int kind = a.gfk[k];
  float coef[LX_MAX_C], inv_den[LX_MAX_C];  // w_ck / D_c and 1 / D_c (mean), w_ck and 1 (sum)
#pragma unroll
  for (int c = 0; c < LX_MAX_C; ++c) {
    float den = 0.f, wck = 0.f;
    if (c < a.C) {
      for (int kk = 0; kk < a.K; ++kk) den = add_nc(den, pw[c * a.K + kk]);
      wck = pw[c * a.K + k];
    }
    inv_den[c] = a.average ? 1.f / den : 1.f;
    coef[c] = a.average ? wck / den : wck;
  }
  float acc[2 + LX_MAX_C];
#pragma unroll
  for (int i = 0; i < 2 + LX_MAX_C; ++i) acc[i] = 0.f;
  const float* Uk = a.U + bk * a.cells;
  float* dUk = a.dU + bk * a.cells;
  for (int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; p < a.cells; p += int64_t(gridDim.x) * blockDim.x) {
    const float u = Uk[p];
    float dgdu, dgdm, dgds;
    // the value as the forward passes compute it (growth_rcp), so that g - G cancels exactly
    // where the field is one kernel's growth
    const float g = growth_rcp(kind, gm, gs, inv_s, inv_9s2, u);
    growth_grad(kind, gm, gs, inv_s, inv_9s2, u, dgdu, dgdm, dgds);
    float dg = 0.f;
#pragma unroll
    for (int c = 0; c < LX_MAX_C; ++c) {
      if (c >= a.C) break;
      const int64_t o = (b * a.C + c) * a.cells + p;
      const float Gc = a.G[o];
      float dGc = dtw * adj_dz(a, a.A[o], Gc, a.lam[o], dtw);
      if (a.dfield) dGc += a.dfield[o];
      dg += dGc * coef[c];
      acc[2 + c] += dGc * (a.average ? (g - Gc) * inv_den[c] : g);
    }
    dUk[p] = a.dpot_in ? fmaf(dg, dgdu, a.dpot_in[bk * a.cells + p]) : dg * dgdu;
    acc[0] += dg * dgdm;
    acc[1] += dg * dgds;
  }
  block_partial<2 + LX_MAX_C>(acc, a.part_g);
}

// One channel, one kernel: the whole pointwise reverse in one pass (grid: blocks per image x B).
// Reads the potential and lambda, and only what the rule needs beside them: the state for v2 and
// the true clip, the field when the caller kept it -- without it, the field is the kernel's growth
// recomputed from the potential as the forward computes it, (w g) / w (mean) or w g (sum).  lambda
// is not rewritten when the step passes it through unchanged (straight-through v1 and simple, in
// place).  lambda_out may alias lambda (each thread reads a cell before it writes it).  KIND is the
// kernel's growth function, so that its value and derivative are selected at compile time instead
// of per cell (same expressions as growth_rcp / growth_grad: the same values).
template <int KIND>
__global__ __launch_bounds__(256) void adj_11_kernel(AdjArgs a) {
  const int64_t b = blockIdx.y;
  const float w = a.w[b * a.w_bs];
  const float gm = a.gf[b * a.gf_bs], gs = a.gf[b * a.gf_bs + 1];
  const float inv_s = 1.f / gs, inv_9s2 = 1.f / (9.f * (gs * gs));
  const float dtw = a.dt[b * a.dt_bs];
  constexpr int kind = KIND;
  const float coef = a.average ? w / w : w;      // dG -> dg (weighted mean of one kernel: w / w)
  const float inv_den = a.average ? 1.f / w : 1.f;
  const bool needA = a.state_fn == LX_STATE_V2 || (a.state_fn == LX_STATE_V1 && !a.ste);
  const bool haveG = a.G != nullptr;
  const bool passes = a.state_fn == LX_STATE_SIMPLE || (a.state_fn == LX_STATE_V1 && a.ste);
  const bool writeLam = !(passes && a.lam_out == a.lam);
  const int64_t base = b * a.cells;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};  // dm, ds, dw, ddt
  // NB cells per round, all loaded before any store: the stores may alias the loads (lambda_out
  // is lambda) only at the same cell, so the next cells' loads need not wait for them.  Cells
  // are taken in the same order as one at a time (the sums round the same way).
  constexpr int NB = 4;
  const int64_t S = int64_t(gridDim.x) * blockDim.x;
  for (int64_t p0 = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; p0 < a.cells; p0 += NB * S) {
    float vA[NB], vG[NB], vU[NB], vL[NB], vF[NB], vP[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int64_t p = p0 + u * S;
      const int64_t q = base + (p < a.cells ? p : p0);  // an in-bounds cell for the tail
      vA[u] = needA ? a.A[q] : 0.f;
      vG[u] = haveG ? a.G[q] : 0.f;
      vU[u] = a.U[q]; vL[u] = a.lam[q];
      vF[u] = a.dfield ? a.dfield[q] : 0.f;
      vP[u] = a.dpot_in ? a.dpot_in[q] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int64_t p = p0 + u * S;
      if (p >= a.cells) break;
      const float A = vA[u], U = vU[u];
      float dgdu, dgdm, dgds;
      const float g = growth_rcp(kind, gm, gs, inv_s, inv_9s2, U);  // as the forward passes
      growth_grad(kind, gm, gs, inv_s, inv_9s2, U, dgdu, dgdm, dgds);
      const float wg = mul_nc(w, g);
      const float G = haveG ? vG[u] : (a.average ? wg / w : wg);
      const float dz = adj_dz(a, A, G, vL[u], dtw);
      float dG = dtw * dz;
      if (a.dfield) dG += vF[u];
      const float dg = dG * coef;
      a.dU[base + p] = a.dpot_in ? fmaf(dg, dgdu, vP[u]) : dg * dgdu;
      acc[0] += dg * dgdm;
      acc[1] += dg * dgds;
      acc[2] += dG * (a.average ? (g - G) * inv_den : g);
      if (a.state_fn == LX_STATE_V2) {
        a.lam_out[base + p] = dz * (1.f - dtw);
        acc[3] += dz * (G - A);
      } else {
        if (writeLam) a.lam_out[base + p] = dz;
        acc[3] += dz * G;
      }
    }
  }
  block_partial<4>(acc, a.part_g);
}

// grid (blocks per image, B * C): direct part of dL/dA and dL/ddt
__global__ __launch_bounds__(256) void adj_state_kernel(AdjArgs a) {
  const int64_t bc = blockIdx.y;
  const int64_t b = bc / a.C;
  const float dtw = a.dt[b * a.dt_bs];
  const int64_t base = bc * a.cells;
  float acc[1] = {0.f};
  for (int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; p < a.cells; p += int64_t(gridDim.x) * blockDim.x) {
    const float A = a.A[base + p], G = a.G[base + p];
    const float dz = adj_dz(a, A, G, a.lam[base + p], dtw);
    if (a.state_fn == LX_STATE_V2) {
      a.lam_out[base + p] = dz * (1.f - dtw);
      acc[0] += dz * (G - A);
    } else {
      a.lam_out[base + p] = dz;
      acc[0] += dz * G;
    }
  }
  block_partial<1>(acc, a.part_s);
}

// The gradient scalars from the blocks' partial sums, in a fixed order: first each image's sum over
// its blocks (ysum[y][i], y the grid row), then each gradient element adds the sums of the worlds
// and channels that feed it (every element has one writer).  Deterministic, and the same for a
// world whatever the batch it is in: the blocks per image depend on the world size alone.
__global__ void adj_ysum_kernel(const float* __restrict__ part, float* __restrict__ ysum, int64_t Y, int NV,
                                int per_img) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= Y * NV) return;
  const int64_t y = i / NV;
  const int s = int(i - y * NV);
  float acc = 0.f;
  for (int blk = 0; blk < per_img; ++blk) acc += part[(y * per_img + blk) * NV + s];
  ysum[i] = acc;
}

__global__ void adj_targets_kernel(AdjArgs a, int64_t B, int one) {
  const int C = a.C, K = a.K;
  const int64_t ngf = a.g_gf ? (a.gf_bs ? B : 1) * 2 * K : 0;
  const int64_t nw = a.g_w ? (a.w_bs ? B : 1) * int64_t(C) * K : 0;
  const int64_t ndt = a.g_dt ? (a.dt_bs ? B : 1) : 0;
  const int NG = 2 + LX_MAX_C;  // slots of a growth row (dm, ds, dw per channel), as block_partial writes it
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < ngf + nw + ndt;
       i += int64_t(gridDim.x) * blockDim.x) {
    float acc = 0.f;
    if (i < ngf) {
      const int64_t bt = i / (2 * K);
      const int e = int(i - bt * 2 * K), k = e / 2, j = e % 2;
      const int64_t b0 = a.gf_bs ? bt : 0, b1 = a.gf_bs ? bt + 1 : B;
      for (int64_t b = b0; b < b1; ++b) acc += one ? a.ysum_g[b * 4 + j] : a.ysum_g[(b * K + k) * NG + j];
      a.g_gf[bt * a.gf_bs + e] += acc;
    } else if (i < ngf + nw) {
      const int64_t r = i - ngf;
      const int64_t bt = r / (int64_t(C) * K);
      const int e = int(r - bt * C * K), c = e / K, k = e % K;
      const int64_t b0 = a.w_bs ? bt : 0, b1 = a.w_bs ? bt + 1 : B;
      for (int64_t b = b0; b < b1; ++b) acc += one ? a.ysum_g[b * 4 + 2] : a.ysum_g[(b * K + k) * NG + 2 + c];
      a.g_w[bt * a.w_bs + e] += acc;
    } else {
      const int64_t bt = i - ngf - nw;
      const int64_t b0 = a.dt_bs ? bt : 0, b1 = a.dt_bs ? bt + 1 : B;
      for (int64_t b = b0; b < b1; ++b) {
        if (one) {
          acc += a.ysum_g[b * 4 + 3];
        } else {
          for (int c = 0; c < C; ++c) acc += a.ysum_s[b * C + c];
        }
      }
      a.g_dt[bt * a.dt_bs] += acc;
    }
  }
}

// blocks per image of the pointwise adjoint: about 4 cells per thread, at most 256
unsigned adj_blocks_per_image(int64_t cells) {
  return unsigned(std::min<int64_t>(std::max<int64_t>((cells + 1023) / 1024, 1), 256));
}

// partial-sum rows per image the workspace holds: the pointwise kernels' blocks, and for one channel
// and one kernel the blocks of lx_adjoint_step's fused row pass (a row pair per lane group)
int64_t adj_rows_per_image(const lx_rule* r) {
  int64_t per = adj_blocks_per_image(cells_of(r));
  if (r->C == 1 && r->K == 1) {
    const int W = r->dims[r->ndim - 1];
    const int64_t rows = cells_of(r) / W;
    per = std::max<int64_t>(per, (rows / 2 + 256 / lx_row_tpr(W) - 1) / (256 / lx_row_tpr(W)));
  }
  return per;
}

// partial-sum layout in the workspace: growth (or one-kernel) rows, state rows, then their sums
struct AdjWork {
  size_t part_g, part_s, ysum_g, ysum_s, total;
};
AdjWork adj_work_layout(const lx_rule* r, int64_t B) {
  const int64_t per = adj_rows_per_image(r);
  const bool one = r->C == 1 && r->K == 1;
  const int64_t yg = one ? B : B * r->K, ys = one ? 0 : B * r->C;
  const int ng = one ? 4 : 2 + LX_MAX_C, ns = one ? 0 : 1;  // row widths of block_partial<4> / <2 + LX_MAX_C>
  AdjWork l;
  size_t o = 0;
  l.part_g = o; o += size_t(yg * per * ng) * sizeof(float);
  l.part_s = o; o += size_t(ys * per * ns) * sizeof(float);
  l.ysum_g = o; o += size_t(yg * ng) * sizeof(float);
  l.ysum_s = o; o += size_t(ys * ns) * sizeof(float);
  l.total = (o + 255) / 256 * 256;
  return l;
}

}  // namespace

// the partial-sum rows of the fused one-kernel step inside an lx_adjoint_workspace_size workspace
float* lx_adj_work_11(const lx_rule* r, int64_t B, void* work) {
  return reinterpret_cast<float*>(static_cast<char*>(work) + adj_work_layout(r, B).part_g);
}

extern "C" int lx_adjoint_workspace_size(const lx_rule* rule, int64_t B, size_t* bytes) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!bytes || B < 0) return lx_fail(LX_EINVAL, "bad arguments");
  *bytes = adj_work_layout(rule, B).total;
  return LX_OK;
}

extern "C" int lx_adjoint_pointwise(const lx_rule* rule, int64_t B, const float* state_d, const float* field_d,
                                    const float* potential_d, const float* gf_d, int64_t gf_bstride, const float* w_d,
                                    int64_t w_bstride, const float* dt_d, int64_t dt_bstride, const float* lambda_d,
                                    const float* dfield_d, const float* dpot_in_d, int32_t ste, float* lambda_out_d,
                                    float* dpot_d, float* grad_gf_d, float* grad_w_d, float* grad_dt_d, void* work_d,
                                    void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  const bool one = rule->C == 1 && rule->K == 1;
  if (!state_d || (!field_d && !one) || !potential_d || !gf_d || !w_d || !dt_d || !lambda_d || !lambda_out_d ||
      !dpot_d || !work_d)
    return lx_fail(LX_EINVAL, "NULL pointer");
  if (B * std::max(rule->C, rule->K) > 65535) return lx_fail(LX_ENOTSUP, "lx_adjoint_pointwise: B * max(C, K) > 65535");
  AdjArgs a;
  a.A = state_d; a.G = field_d; a.U = potential_d; a.lam = lambda_d; a.dfield = dfield_d; a.dpot_in = dpot_in_d;
  a.gf = gf_d; a.gf_bs = gf_bstride; a.w = w_d; a.w_bs = w_bstride; a.dt = dt_d; a.dt_bs = dt_bstride;
  a.lam_out = lambda_out_d; a.dU = dpot_d; a.g_gf = grad_gf_d; a.g_w = grad_w_d; a.g_dt = grad_dt_d;
  a.cells = cells_of(rule); a.C = rule->C; a.K = rule->K;
  for (int k = 0; k < LX_MAX_K; ++k) {
    a.cin[k] = k < rule->K ? rule->cin[k] : 0;
    a.gfk[k] = k < rule->K ? rule->gf[k] : 0;
  }
  a.state_fn = rule->state_fn; a.average = rule->average; a.ste = ste != 0;
  const AdjWork l = adj_work_layout(rule, B);
  char* wk = static_cast<char*>(work_d);
  a.part_g = reinterpret_cast<float*>(wk + l.part_g);
  a.part_s = reinterpret_cast<float*>(wk + l.part_s);
  a.ysum_g = reinterpret_cast<float*>(wk + l.ysum_g);
  a.ysum_s = reinterpret_cast<float*>(wk + l.ysum_s);
  // blocks per image from the world size alone: each block leaves its partial gradient sums in the
  // workspace (no atomics), summed afterwards in a fixed order
  const unsigned per_img = adj_blocks_per_image(a.cells);
  hipStream_t st = (hipStream_t)stream;
  if (one) {
    const dim3 grid(per_img, unsigned(B));
    switch (rule->gf[0]) {
      case LX_GF_POLY_QUAD4: hipLaunchKernelGGL(adj_11_kernel<LX_GF_POLY_QUAD4>, grid, dim3(256), 0, st, a); break;
      case LX_GF_GAUSSIAN: hipLaunchKernelGGL(adj_11_kernel<LX_GF_GAUSSIAN>, grid, dim3(256), 0, st, a); break;
      case LX_GF_GAUSSIAN_TARGET:
        hipLaunchKernelGGL(adj_11_kernel<LX_GF_GAUSSIAN_TARGET>, grid, dim3(256), 0, st, a);
        break;
      case LX_GF_STEP: hipLaunchKernelGGL(adj_11_kernel<LX_GF_STEP>, grid, dim3(256), 0, st, a); break;
      case LX_GF_STAIRCASE: hipLaunchKernelGGL(adj_11_kernel<LX_GF_STAIRCASE>, grid, dim3(256), 0, st, a); break;
      case LX_GF_TRIANGLE: hipLaunchKernelGGL(adj_11_kernel<LX_GF_TRIANGLE>, grid, dim3(256), 0, st, a); break;
      default: hipLaunchKernelGGL(adj_11_kernel<LX_GF_IDENTITY>, grid, dim3(256), 0, st, a); break;
    }
  } else {
    // growth first: it reads lambda, which the state kernel may overwrite in place
    hipLaunchKernelGGL(adj_growth_kernel, dim3(per_img, unsigned(B * rule->K)), dim3(256), 0, st, a);
    hipLaunchKernelGGL(adj_state_kernel, dim3(per_img, unsigned(B * rule->C)), dim3(256), 0, st, a);
  }
  if (grad_gf_d || grad_w_d || grad_dt_d) {
    const int64_t yg = one ? B : B * rule->K;
    const int ng = one ? 4 : 2 + LX_MAX_C;
    hipLaunchKernelGGL(adj_ysum_kernel, dim3(grid_for(yg * ng, 256)), dim3(256), 0, st, a.part_g, a.ysum_g, yg, ng,
                       int(per_img));
    if (!one)
      hipLaunchKernelGGL(adj_ysum_kernel, dim3(grid_for(B * rule->C, 256)), dim3(256), 0, st, a.part_s, a.ysum_s,
                         B * rule->C, 1, int(per_img));
    hipLaunchKernelGGL(adj_targets_kernel, dim3(1), dim3(256), 0, st, a, B, one ? 1 : 0);
  }
  return launched("adjoint pointwise");
}

extern "C" int lx_stats(const lx_rule* rule, const float* cells_d, const float* field_d, const float* potential_d,
                        int64_t B, float R, float T, void* carry_d, float* stats_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (B < 0 || !(R > 0.f) || !(T > 0.f)) return lx_fail(LX_EINVAL, "bad B / R / T");
  if (B == 0) return LX_OK;
  if (!cells_d || !field_d || !potential_d || !carry_d || !stats_d) return lx_fail(LX_EINVAL, "NULL pointer");
  StatArgs sa;
  sa.cells = cells_d; sa.field = field_d; sa.pot = potential_d;
  sa.carry = static_cast<lx_stat_carry*>(carry_d);
  sa.stats = stats_d; sa.B = B; sa.ndim = rule->ndim;
  for (int i = 0; i < 3; ++i) sa.dims[i] = i < rule->ndim ? rule->dims[i] : 1;
  sa.C = rule->C; sa.K = rule->K;
  sa.R2 = (float)((double)R * (double)R); sa.R = R; sa.dt = (float)(1.0 / (double)T);
  if (B > 65535) return lx_fail(LX_ENOTSUP, "lx_stats: more than 65535 worlds per call");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(stats_part_kernel, dim3(LX_STAT_PARTS, (unsigned)B), dim3(256), 0, st, sa);
  hipLaunchKernelGGL(stats_finish_kernel, dim3(grid_for(B, 64)), dim3(64), 0, st, sa);
  return launched("stats");
}

int lx_pair_stats_launch(const lx_rule* rule, const float* part, int64_t B, int64_t pairs, float R, float T,
                         void* carry_d, float* stats_d, hipStream_t st) {
  StatArgs sa;
  sa.cells = nullptr; sa.field = nullptr; sa.pot = nullptr;
  sa.carry = static_cast<lx_stat_carry*>(carry_d);
  sa.stats = stats_d; sa.B = B; sa.ndim = rule->ndim;
  for (int i = 0; i < 3; ++i) sa.dims[i] = i < rule->ndim ? rule->dims[i] : 1;
  sa.C = rule->C; sa.K = rule->K;
  sa.R2 = (float)((double)R * (double)R); sa.R = R; sa.dt = (float)(1.0 / (double)T);
  if (B > 2147483647) return lx_fail(LX_ENOTSUP, "too many worlds");
  hipLaunchKernelGGL(pair_stats_kernel, dim3((unsigned)B), dim3(kPairThreads), 0, st, sa, part, pairs);
  return launched("pair stats");
}
