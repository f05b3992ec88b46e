// Plain (unfused) path: the leniax step as three HBM passes, plus the elementwise
// field / state kernels, the per-world statistics and the heuristic scan.
//
//   rows_fwd:  real rows -> half spectra along W (two rows packed per complex FFT)
//   cols:      FFT along H of every half-spectrum column, multiply by each kernel's
//              half spectrum (depthwise: kernel k reads channel cin[k]), inverse FFT
//   rows_inv:  per channel c: for every kernel k feeding c, inverse FFT along W of
//              two rows at once, growth g_k, weighted sum into G_c; then the update
// Workspace: S [B][C][H][W/2+1] and U^ [B][K][H][W/2+1] complex64.
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
    for (int k = 0; k < rd.K; ++k) g[k] = growth(rd.gf[k], pg[2 * k], pg[2 * k + 1], pot[(b * rd.K + k) * n + cell]);
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

__global__ void carry_init_kernel(lx_carry* c, int64_t B) {
  const int64_t b = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (b >= B) return;
  lx_carry z;
  for (int i = 0; i < 3; ++i) { z.shift[i] = 0; z.centroid[i] = 0.f; }
  z.angle = 0.f;
  for (int i = 0; i < LX_MAX_C; ++i) z.init_cm[i] = 0.f;
  z.prev_mass = z.prev_sign = 0.f;
  z.mono = z.vol = z.n = z.t = z.pad = 0;
  z.cont = 1;
  c[b] = z;
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
  if (!potential_d || !gf_d || !w_d || !field_d) return lx_fail(LX_EINVAL, "NULL pointer");
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  const int64_t n = cells_of(rule);
  hipLaunchKernelGGL(field_kernel, dim3(grid_for(B * n, 256)), dim3(256), 0, (hipStream_t)stream, to_dev(rule),
                     potential_d, B, n, gf_d, gf_bstride, w_d, w_bstride, field_d);
  return launched("field");
}

extern "C" int lx_state(const lx_rule* rule, const float* state_d, const float* field_d, int64_t B, const float* dt_d,
                        int64_t dt_bstride, float* out_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!state_d || !field_d || !dt_d || !out_d) return lx_fail(LX_EINVAL, "NULL pointer");
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  const int64_t nc = cells_of(rule) * rule->C;
  hipLaunchKernelGGL(state_kernel, dim3(grid_for(B * nc, 256)), dim3(256), 0, (hipStream_t)stream, rule->state_fn,
                     state_d, field_d, B, nc, dt_d, dt_bstride, out_d);
  return launched("state");
}

extern "C" int lx_check_heuristics(const float* stats_d, int32_t T, int64_t B, int32_t C, float* cont_d, float* n_d,
                                   void* stream) {
  if (!stats_d) return lx_fail(LX_EINVAL, "NULL stats");
  if (T < 1 || B < 0 || C < 1 || C > LX_MAX_C) return lx_fail(LX_EINVAL, "bad T/B/C");
  if (B == 0) return LX_OK;
  hipLaunchKernelGGL(heuristics_kernel, dim3(grid_for(B, 128)), dim3(128), 0, (hipStream_t)stream, stats_d, T, B, C,
                     cont_d, n_d);
  return launched("check_heuristics");
}

extern "C" size_t lx_carry_bytes(void) { return sizeof(lx_carry); }

extern "C" int lx_carry_init(int64_t B, void* carry_d, void* stream) {
  if (!carry_d || B < 0) return lx_fail(LX_EINVAL, "bad carry arguments");
  if (B == 0) return LX_OK;
  hipLaunchKernelGGL(carry_init_kernel, dim3(grid_for(B, 128)), dim3(128), 0, (hipStream_t)stream,
                     static_cast<lx_carry*>(carry_d), B);
  return launched("carry_init");
}


// ====================================================================== FFT path
namespace {

template <int L>
struct Lanes {
  static constexpr int TPR = L <= 128 ? 4 : L / 32;  // lanes per transform
  static constexpr int P = L / TPR;                   // points per lane
};

// twiddle table exp(-2 pi i j / L) in global memory, one per supported length
template <int L>
__device__ float2 g_tw[L];

template <int L>
__global__ void init_tw_kernel() {
  for (int j = threadIdx.x; j < L; j += blockDim.x) {
    double sn, cs;
    sincospi(2.0 * j / L, &sn, &cs);
    g_tw<L>[j] = make_float2((float)cs, (float)-sn);
  }
}

// partner lane holding block (TPR - s) mod TPR, s = bitrev(t)
template <int TPR>
LX_DEV int neg_block_lane(int t) {
  const int s = bitrev_n(t, TPR);
  return bitrev_n((TPR - s) & (TPR - 1), TPR);
}

// value of lane t ^ (TPR-1) in the same lane group
template <int TPR>
LX_DEV float2 mirror_fetch(float2 v) {
  if constexpr (TPR == 4) return dpp2<LX_QP(3, 2, 1, 0)>(v);
  else if constexpr (TPR == 8) return dpp2<0x141>(v);   // row_half_mirror
  else return dpp2<0x140>(v);                           // row_mirror (TPR == 16)
}

// ---------------------------------------------------------------- rows forward
template <int W>
__global__ __launch_bounds__(256) void rows_fwd_kernel(const float* __restrict__ A, float2* __restrict__ S,
                                                       int64_t npairs, int H) {
  using LC = Lanes<W>;
  constexpr int TPR = LC::TPR, P = LC::P, WH = W / 2 + 1;
  const int lane = threadIdx.x & 63;
  const int t = threadIdx.x % TPR;
  const int64_t pr = (int64_t(blockIdx.x) * 256 + threadIdx.x) / TPR;
  const bool live = pr < npairs;
  const int64_t img = live ? pr / (H / 2) : 0;
  const int r2 = live ? int(pr % (H / 2)) * 2 : 0;
  const float* a0 = A + (img * H + r2) * W;
  const float* a1 = a0 + W;
  float2 x[P];
#pragma unroll
  for (int m = 0; m < P; ++m) x[m] = make_float2(a0[TPR * m + t], a1[TPR * m + t]);
  fft_lanes<P, TPR, false>(x, t, lane, g_tw<W>);
  const int s = bitrev_n(t, TPR);
  // Z[W - k]: k1 >= 1 from lane t^(TPR-1) at index P - k1; k1 == 0 from block (TPR - s) at index 0
  const float2 z0n = make_float2(__shfl(x[0].x, (lane & ~(TPR - 1)) | neg_block_lane<TPR>(t), 64),
                                 __shfl(x[0].y, (lane & ~(TPR - 1)) | neg_block_lane<TPR>(t), 64));
  float2* s0 = S + (img * H + r2) * int64_t(WH);
  float2* s1 = s0 + WH;
#pragma unroll
  for (int k1 = 0; k1 < P; ++k1) {
    const float2 zn = k1 == 0 ? z0n : mirror_fetch<TPR>(x[P - k1 > 0 && k1 > 0 ? P - k1 : 0]);
    const float2 z = x[k1];
    const int k = k1 + P * s;
    // X_a = (z + conj zn) / 2, X_b = (z - conj zn) / 2i
    const float2 xa = make_float2(0.5f * (z.x + zn.x), 0.5f * (z.y - zn.y));
    const float2 xb = make_float2(0.5f * (z.y + zn.y), 0.5f * (zn.x - z.x));
    if (live && k <= W / 2) {
      s0[k] = xa;
      s1[k] = xb;
    }
  }
}

// ---------------------------------------------------------------- columns
struct ColArgs {
  const float2* S;
  float2* U;
  const float2* Kh;
  int64_t k_bstride;
  int64_t ncols;  // B * C * WH
  int C, K;
  int cin[LX_MAX_K];
  int mode;    // 0: conv (forward, x kernel spectra, inverse), 1: forward only (in place), 2: inverse only S -> U
  int conjk;   // multiply by conj(K) (adjoint of the potential)
  float2* FA;  // optional forward spectra output [img][H][WH] (mode 0)
  // optional kernel-spectrum gradient: dspec[b][k] += X_k * conj(FAin[b][acc_cin[k]])
  const float2* FAin;
  float2* dspec;
  int C_fa;
  int acc_cin[LX_MAX_K];
};

ColArgs col_args() {
  ColArgs a;
  a.S = nullptr; a.U = nullptr; a.Kh = nullptr; a.k_bstride = 0; a.ncols = 0; a.C = 1; a.K = 0;
  for (int k = 0; k < LX_MAX_K; ++k) { a.cin[k] = -1; a.acc_cin[k] = 0; }
  a.mode = 0; a.conjk = 0; a.FA = nullptr; a.FAin = nullptr; a.dspec = nullptr; a.C_fa = 1;
  return a;
}

template <int H, int W>
__global__ __launch_bounds__(256) void cols_kernel(ColArgs a) {
  using LC = Lanes<H>;
  constexpr int TPR = LC::TPR, P = LC::P, WH = W / 2 + 1;
  const int lane = threadIdx.x & 63;
  const int t = threadIdx.x % TPR;
  const int64_t cid = (int64_t(blockIdx.x) * 256 + threadIdx.x) / TPR;
  const bool live = cid < a.ncols;
  const int64_t img = live ? cid / WH : 0;
  const int j = live ? int(cid % WH) : 0;
  const int64_t b = img / a.C;
  const int c = int(img % a.C);
  const float2* src = a.S + img * int64_t(H) * WH + j;
  float2 x[P], y[P];
  if (a.mode == 2) {  // inverse transform only (input in frequency order), S -> U
    const int s = bitrev_n(t, TPR);
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) y[k1] = src[int64_t(k1 + P * s) * WH];
    fft_lanes<P, TPR, true>(y, t, lane, g_tw<H>);
    float2* dst = a.U + img * int64_t(H) * WH + j;
    if (live) {
#pragma unroll
      for (int m = 0; m < P; ++m) dst[int64_t(TPR * m + t) * WH] = y[m];
    }
    return;
  }
#pragma unroll
  for (int m = 0; m < P; ++m) x[m] = src[int64_t(TPR * m + t) * WH];
  fft_lanes<P, TPR, false>(x, t, lane, g_tw<H>);
  const int s = bitrev_n(t, TPR);
  if (a.mode == 1) {  // forward transform only, in place (spectra of real arrays)
    float2* dst = const_cast<float2*>(a.S) + img * int64_t(H) * WH + j;
    if (live) {
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) dst[int64_t(k1 + P * s) * WH] = x[k1];
    }
    return;
  }
  if (a.FA && live) {
    float2* fa = a.FA + img * int64_t(H) * WH + j;
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) fa[int64_t(k1 + P * s) * WH] = x[k1];
  }
  for (int k = 0; k < a.K; ++k) {
    if (a.cin[k] != c) continue;
    if (a.dspec && live) {
      const float2* fa = a.FAin + (b * a.C_fa + a.acc_cin[k]) * int64_t(H) * WH + j;
      float2* ds = a.dspec + (b * a.K + k) * int64_t(H) * WH + j;
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) {
        const int64_t o = int64_t(k1 + P * s) * WH;
        ds[o] = cadd(ds[o], cmulc(x[k1], fa[o]));
      }
    }
    const float2* kk = a.Kh + b * a.k_bstride + int64_t(k) * H * WH + j;
    if (a.conjk) {
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) y[k1] = cmulc(x[k1], kk[int64_t(k1 + P * s) * WH]);
    } else {
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) y[k1] = cmul(x[k1], kk[int64_t(k1 + P * s) * WH]);
    }
    fft_lanes<P, TPR, true>(y, t, lane, g_tw<H>);
    float2* dst = a.U + (b * a.K + k) * int64_t(H) * WH + j;
    if (live) {
#pragma unroll
      for (int m = 0; m < P; ++m) dst[int64_t(TPR * m + t) * WH] = y[m];
    }
  }
}

// ---------------------------------------------------------------- rows inverse + field + update
struct RowArgs {
  const float2* U;
  const float* A;
  float* A_out;
  float* field;
  float* pot;
  const float* gf;
  int64_t gf_bs;
  const float* w;
  int64_t w_bs;
  const float* dt;
  int64_t dt_bs;
  int64_t npairs;  // B * C * H / 2
  int H, C, K;
  int gfk[LX_MAX_K];
  int state_fn, average;
  int update;      // 0: potential only
  // selection mode: weight of kernel k in channel c is (cin[k] == c), growth is the identity and the
  // update is A_out = A + G (no gf / w / dt reads).  Used for the potential alone, the adjoint of the
  // potential (lambda += correlations) and the inverse real FFT.
  int sel;
  int cin[LX_MAX_K];
};

template <int W>
__global__ __launch_bounds__(256) void rows_inv_kernel(RowArgs a) {
  using LC = Lanes<W>;
  constexpr int TPR = LC::TPR, P = LC::P, WH = W / 2 + 1;
  const int lane = threadIdx.x & 63;
  const int t = threadIdx.x % TPR;
  const int64_t pr = (int64_t(blockIdx.x) * 256 + threadIdx.x) / TPR;
  const bool live = pr < a.npairs;
  const int64_t img = live ? pr / (a.H / 2) : 0;
  const int r2 = live ? int(pr % (a.H / 2)) * 2 : 0;
  const int64_t b = img / a.C;
  const int c = int(img % a.C);
  const int64_t HWp = int64_t(a.H) * W;
  const float inv_n = 1.f / float(HWp);
  const float* pw = a.sel ? nullptr : a.w + b * a.w_bs;
  const float* pg = a.sel ? nullptr : a.gf + b * a.gf_bs;
  const int s = bitrev_n(t, TPR);
  const bool mirror = s >= TPR / 2;
  const float sg = mirror ? -1.f : 1.f;
  float G0[P], G1[P];
#pragma unroll
  for (int m = 0; m < P; ++m) G0[m] = G1[m] = 0.f;
  float den = 0.f;
  float2 x[P];
  for (int k = 0; k < a.K; ++k) {
    const float wk = a.sel ? (a.cin[k] == c ? 1.f : 0.f) : pw[c * a.K + k];
    den = add_nc(den, wk);
    int owner = 0;
    if (a.sel) {
      owner = a.cin[k];
    } else {
      for (int cc = a.C - 1; cc >= 0; --cc)
        if (pw[cc * a.K + k] != 0.f) owner = cc;
    }
    const bool want_pot = a.pot != nullptr && owner == c;
    if (wk == 0.f && !want_pot) continue;
    const float2* u0 = a.U + ((b * a.K + k) * int64_t(a.H) + r2) * WH;
    const float2* u1 = u0 + WH;
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      const int kk = k1 + P * s;
      const int col = mirror ? W - kk : kk;  // W - kk <= W/2 when mirrored
      const float2 ya = u0[col], yb = u1[col];
      float2 z = make_float2(fmaf(-sg, yb.y, ya.x), fmaf(sg, ya.y, yb.x));
      if (k1 == 0 && (s == 0 || s == TPR / 2)) z = make_float2(ya.x, yb.x);  // DC and Nyquist are real
      x[k1] = z;
    }
    fft_lanes<P, TPR, true>(x, t, lane, g_tw<W>);
    const float gm = a.sel ? 0.f : pg[2 * k], gs = a.sel ? 0.f : pg[2 * k + 1];
    float* po0 = want_pot ? a.pot + ((b * a.K + k) * int64_t(a.H) + r2) * W : nullptr;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const float U0 = x[m].x * inv_n, U1 = x[m].y * inv_n;
      if (want_pot && live) {
        po0[TPR * m + t] = U0;
        po0[W + TPR * m + t] = U1;
      }
      if (a.update) {
        G0[m] = add_nc(G0[m], mul_nc(wk, a.sel ? U0 : growth(a.gfk[k], gm, gs, U0)));
        G1[m] = add_nc(G1[m], mul_nc(wk, a.sel ? U1 : growth(a.gfk[k], gm, gs, U1)));
      }
    }
  }
  if (!a.update || !live) return;
  const float dtw = a.sel ? 1.f : a.dt[b * a.dt_bs];
  const int64_t off = (img * a.H + r2) * int64_t(W);
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int col = TPR * m + t;
    float g0 = G0[m], g1 = G1[m];
    if (a.average) { g0 = (g0 / den); g1 = (g1 / den); }
    if (a.field) { a.field[off + col] = g0; a.field[off + W + col] = g1; }
    const float v0 = a.A[off + col], v1 = a.A[off + W + col];
    float n0, n1;
    if (a.state_fn == LX_STATE_V2) {
      n0 = add_nc(mul_nc(v0, sub_nc(1.f, dtw)), mul_nc(dtw, g0));
      n1 = add_nc(mul_nc(v1, sub_nc(1.f, dtw)), mul_nc(dtw, g1));
    } else {
      n0 = add_nc(v0, mul_nc(dtw, g0));
      n1 = add_nc(v1, mul_nc(dtw, g1));
      if (a.state_fn == LX_STATE_V1) {
        n0 = n0 < 0.f ? 0.f : (n0 > 1.f ? 1.f : n0);
        n1 = n1 < 0.f ? 0.f : (n1 > 1.f ? 1.f : n1);
      }
    }
    a.A_out[off + col] = n0;
    a.A_out[off + W + col] = n1;
  }
}

// ---------------------------------------------------------------- statistics (one block per world)
struct StatArgs {
  const float* cells;
  const float* field;
  const float* pot;
  lx_carry* carry;
  float* stats;
  int64_t B;
  int ndim;
  int dims[3];
  int C, K;
  float R2, R, dt;
};

__global__ __launch_bounds__(256) void stats_kernel(StatArgs a) {
  __shared__ int shs[3];
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  lx_carry c;
  if (tid == 0) {
    c = a.carry[b];
    for (int i = 0; i < 3; ++i) shs[i] = c.shift[i];
  }
  __syncthreads();
  const int d0 = a.ndim == 3 ? a.dims[0] : 1, d1 = a.dims[a.ndim - 2], d2 = a.dims[a.ndim - 1];
  const int64_t n = int64_t(d0) * d1 * d2;
  // ndim == 2: axes (d1, d2) map to centroid components (0, 1); ndim == 3: (d0, d1, d2) -> (0, 1, 2)
  const int s0 = a.ndim == 3 ? shs[0] : 0, s1 = shs[a.ndim - 2], s2 = shs[a.ndim - 1];
  float v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = 0.f;
  // v: 0 sA, 1 sG, 2 nA, 3 nG, 4 nU, 5..7 sAy, 8..10 sGy, 11..13 sAyy (components in world-axis order)
  float cm[LX_MAX_C];
  for (int ch = 0; ch < LX_MAX_C; ++ch) cm[ch] = 0.f;
  for (int ch = 0; ch < a.C; ++ch) {
    const float* A = a.cells + (b * a.C + ch) * n;
    const float* G = a.field + (b * a.C + ch) * n;
    float sc = 0.f;
    for (int64_t i = tid; i < n; i += 256) {
      const int z = int(i % d2), y = int((i / d2) % d1), xq = int(i / (int64_t(d1) * d2));
      const float cz = float(((z - s2) % d2 + d2) % d2 - d2 / 2);
      const float cy = float(((y - s1) % d1 + d1) % d1 - d1 / 2);
      const float cx = float(((xq - s0) % d0 + d0) % d0 - d0 / 2);
      const float av = A[i];
      const float g = G[i];
      const float gp = (g != g) ? g : fmaxf(g, 0.f);
      float co[3];
      if (a.ndim == 3) { co[0] = cx; co[1] = cy; co[2] = cz; } else { co[0] = cy; co[1] = cz; co[2] = 0.f; }
      sc += av;
      v[1] += gp;
      v[2] += av > LX_EPS ? 1.f : 0.f;
      v[3] += gp > LX_EPS ? 1.f : 0.f;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        v[5 + d] = fmaf(av, co[d], v[5 + d]);
        v[8 + d] = fmaf(gp, co[d], v[8 + d]);
        v[11 + d] = fmaf(av * co[d], co[d], v[11 + d]);
      }
    }
    v[0] += sc;
    cm[ch] = sc;
  }
  for (int k = 0; k < a.K; ++k) {
    const float* U = a.pot + (b * a.K + k) * n;
    for (int64_t i = tid; i < n; i += 256) v[4] += U[i] > LX_EPS ? 1.f : 0.f;
  }
  // block reduction of the 14 sums and the C channel sums
  constexpr int NV = 14 + LX_MAX_C;
  __shared__ float red2[4][NV];
  for (int i = 0; i < NV; ++i) {
    const float val = wave_sum(i < 14 ? v[i] : cm[i - 14]);
    if (lane == 0) red2[wave][i] = val;
  }
  __syncthreads();
  float out[NV];
  if (tid == 0)
    for (int i = 0; i < NV; ++i) out[i] = red2[0][i] + red2[1][i] + red2[2][i] + red2[3][i];
  if (tid == 0) {
    lx_sums s;
    s.sA = out[0]; s.sG = out[1]; s.nA = out[2]; s.nG = out[3]; s.nU = out[4];
    for (int d = 0; d < 3; ++d) { s.sAy[d] = out[5 + d]; s.sGy[d] = out[8 + d]; s.sAyy[d] = out[11 + d]; }
    for (int ch = 0; ch < LX_MAX_C; ++ch) s.sAc[ch] = ch < a.C ? out[14 + ch] : 0.f;
    int32_t dd[3];
    for (int i = 0; i < 3; ++i) dd[i] = a.dims[i];
    lx_stats_finish(s, c, a.ndim, dd, a.C, a.R2, a.R, a.dt, LX_STOP_NONE, 0, float(n * a.C),
                    a.stats + b * (LX_NSTAT_BASE + a.C));
    a.carry[b] = c;
  }
}

template <int L>
bool tw_ready_flag = false;

template <int L>
int ensure_tw(hipStream_t stream) {
  if (!tw_ready_flag<L>) {
    hipLaunchKernelGGL(init_tw_kernel<L>, dim3(1), dim3(256), 0, stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lx_fail(LX_EDEVICE, "twiddle init: %s", hipGetErrorString(e));
    tw_ready_flag<L> = true;
  }
  return LX_OK;
}

bool plain_size_ok(int n) { return n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512; }

template <int V>
using IntC = std::integral_constant<int, V>;

// calls f(IntC<H>, IntC<W>) for the supported power-of-two sizes
template <typename F>
int dispatch_hw(int H, int W, F&& f) {
#define LX_W(HH, WW) \
  case WW: return f(IntC<HH>{}, IntC<WW>{});
#define LX_H(HH)                                                                              \
  case HH:                                                                                    \
    switch (W) {                                                                              \
      LX_W(HH, 16) LX_W(HH, 32) LX_W(HH, 64) LX_W(HH, 128) LX_W(HH, 256) LX_W(HH, 512)        \
      default: return lx_fail(LX_ENOTSUP, "width %d (powers of two in [16, 512])", W);        \
    }
  switch (H) {
    LX_H(16) LX_H(32) LX_H(64) LX_H(128) LX_H(256) LX_H(512)
    default: return lx_fail(LX_ENOTSUP, "height %d (powers of two in [16, 512])", H);
  }
#undef LX_H
#undef LX_W
}

template <int H, int W>
int tw_pair(hipStream_t st) {
  int rc = ensure_tw<H>(st);
  return rc ? rc : ensure_tw<W>(st);
}

// real rows [n][H][W] -> half spectra along W
template <int H, int W>
void launch_rows_fwd(const float* A, float2* S, int64_t n, hipStream_t st) {
  const int64_t npairs = n * int64_t(H / 2);
  const int64_t thr = npairs * Lanes<W>::TPR;
  hipLaunchKernelGGL(rows_fwd_kernel<W>, dim3(unsigned((thr + 255) / 256)), dim3(256), 0, st, A, S, npairs, H);
}

template <int H, int W>
void launch_cols(const ColArgs& ca, hipStream_t st) {
  const int64_t thr = ca.ncols * Lanes<H>::TPR;
  hipLaunchKernelGGL((cols_kernel<H, W>), dim3(unsigned((thr + 255) / 256)), dim3(256), 0, st, ca);
}

template <int H, int W>
void launch_rows_inv(const RowArgs& ra, hipStream_t st) {
  const int64_t thr = ra.npairs * Lanes<W>::TPR;
  hipLaunchKernelGGL(rows_inv_kernel<W>, dim3(unsigned((thr + 255) / 256)), dim3(256), 0, st, ra);
}

RowArgs row_args(const lx_rule* rule, int64_t B, int H) {
  RowArgs ra;
  ra.U = nullptr; ra.A = nullptr; ra.A_out = nullptr; ra.field = nullptr; ra.pot = nullptr;
  ra.gf = nullptr; ra.gf_bs = 0; ra.w = nullptr; ra.w_bs = 0; ra.dt = nullptr; ra.dt_bs = 0;
  ra.npairs = B * rule->C * int64_t(H / 2); ra.H = H; ra.C = rule->C; ra.K = rule->K;
  for (int k = 0; k < LX_MAX_K; ++k) {
    ra.gfk[k] = k < rule->K ? rule->gf[k] : 0;
    ra.cin[k] = k < rule->K ? rule->cin[k] : -1;
  }
  ra.state_fn = rule->state_fn; ra.average = rule->average; ra.update = 0; ra.sel = 0;
  return ra;
}

ColArgs conv_args(const lx_rule* rule, int64_t B, int WH) {
  ColArgs ca = col_args();
  ca.ncols = B * rule->C * int64_t(WH); ca.C = rule->C; ca.K = rule->K;
  for (int k = 0; k < LX_MAX_K; ++k) ca.cin[k] = k < rule->K ? rule->cin[k] : -1;
  return ca;
}

// one leniax step (or, with sel, the potential alone) in three passes over HBM
template <int H, int W>
int run_plain(const lx_rule* rule, const float* A, int64_t B, const float2* Kh, int64_t k_bs, const float* gf,
              int64_t gf_bs, const float* w, int64_t w_bs, const float* dt, int64_t dt_bs, float* A_out,
              float* field, float* pot, void* work, hipStream_t st, int sel) {
  if (int rc = tw_pair<H, W>(st)) return rc;
  constexpr int WH = W / 2 + 1;
  float2* S = static_cast<float2*>(work);
  float2* U = S + B * rule->C * int64_t(H) * WH;
  launch_rows_fwd<H, W>(A, S, B * rule->C, st);
  ColArgs ca = conv_args(rule, B, WH);
  ca.S = S; ca.U = U; ca.Kh = Kh; ca.k_bstride = k_bs;
  launch_cols<H, W>(ca, st);
  RowArgs ra = row_args(rule, B, H);
  ra.U = U; ra.A = A; ra.A_out = A_out; ra.field = field; ra.pot = pot;
  ra.gf = gf; ra.gf_bs = gf_bs; ra.w = w; ra.w_bs = w_bs; ra.dt = dt; ra.dt_bs = dt_bs;
  ra.update = A_out != nullptr; ra.sel = sel;
  launch_rows_inv<H, W>(ra, st);
  return launched("plain step");
}

int plain_step(const lx_rule* rule, const float* A, int64_t B, const void* Kh, int64_t k_bs, const float* gf,
               int64_t gf_bs, const float* w, int64_t w_bs, const float* dt, int64_t dt_bs, float* A_out,
               float* field, float* pot, void* work, hipStream_t st, int sel = 0) {
  if (rule->ndim != 2) return lx_fail(LX_ENOTSUP, "plain step: 3-D worlds are not built yet");
  const float2* K2 = static_cast<const float2*>(Kh);
  return dispatch_hw(rule->dims[0], rule->dims[1], [&](auto h, auto w_) {
    return run_plain<decltype(h)::value, decltype(w_)::value>(rule, A, B, K2, k_bs, gf, gf_bs, w, w_bs, dt, dt_bs,
                                                              A_out, field, pot, work, st, sel);
  });
}

// full spectrum from the half spectrum of a real array (Hermitian extension)
__global__ void expand_half_kernel(const float2* __restrict__ half, float2* __restrict__ full, int64_t n, int H, int W,
                                   float scale) {
  const int WH = W / 2 + 1;
  const int64_t total = n * H * int64_t(W);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int kw = int(i % W);
    const int kh = int((i / W) % H);
    const int64_t img = i / (int64_t(H) * W);
    const float2* hb = half + img * int64_t(H) * WH;
    full[i] = cscale(kw <= W / 2 ? hb[int64_t(kh) * WH + kw] : cconj(hb[int64_t((H - kh) % H) * WH + (W - kw)]), scale);
  }
}

// numpy.fft.rfft2 of n real [H][W] arrays (unnormalised)
template <int H, int W>
int rfft2_2d(int64_t n, const float* x, float2* X, hipStream_t st) {
  if (int rc = tw_pair<H, W>(st)) return rc;
  launch_rows_fwd<H, W>(x, X, n, st);
  ColArgs ca = col_args();
  ca.S = X; ca.ncols = n * int64_t(W / 2 + 1); ca.mode = 1;
  launch_cols<H, W>(ca, st);
  return launched("rfft2");
}

// numpy.fft.irfft2 (1/(H W) scaling) of n half spectra; work holds n H (W/2+1) complex
template <int H, int W>
int irfft2_2d(int64_t n, const float2* X, float* x, float2* work, hipStream_t st) {
  if (int rc = tw_pair<H, W>(st)) return rc;
  ColArgs ca = col_args();
  ca.S = X; ca.U = work; ca.ncols = n * int64_t(W / 2 + 1); ca.mode = 2;
  launch_cols<H, W>(ca, st);
  lx_rule one{};
  one.ndim = 2; one.dims[0] = H; one.dims[1] = W; one.C = 1; one.K = 1; one.cin[0] = 0;
  RowArgs ra = row_args(&one, n, H);
  ra.U = work; ra.pot = x; ra.sel = 1;
  launch_rows_inv<H, W>(ra, st);
  return launched("irfft2");
}

template <int H, int W>
int spectrum_2d(int64_t n, const float* k, float2* full, float2* work, hipStream_t st) {
  if (int rc = rfft2_2d<H, W>(n, k, work, st)) return rc;
  hipLaunchKernelGGL(expand_half_kernel, dim3(grid_for(n * H * int64_t(W), 256)), dim3(256), 0, st, work, full, n, H,
                     W, 1.f);
  return launched("kernel spectrum");
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

// block-wide sums of NV per-thread values, each added atomically into dst[i] (skipped where NULL)
template <int NV>
LX_DEV void block_atomic_sum(float (&v)[NV], float* const (&dst)[NV]) {
  __shared__ float red[NV][4];
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const float s = wave_sum(v[i]);
    if (ln == 0) red[i][wv] = s;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i)
    if (threadIdx.x == i && dst[i]) atomicAdd(dst[i], red[i][0] + red[i][1] + red[i][2] + red[i][3]);
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
    const float g = growth(kind, gm, gs, u);
    float dgdu, dgdm, dgds;
    growth_grad(kind, gm, gs, u, dgdu, dgdm, dgds);
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
  float* const dst[2 + LX_MAX_C] = {
      a.g_gf ? a.g_gf + b * a.gf_bs + 2 * k : nullptr, a.g_gf ? a.g_gf + b * a.gf_bs + 2 * k + 1 : nullptr,
      a.g_w && 0 < a.C ? a.g_w + b * a.w_bs + 0 * a.K + k : nullptr,
      a.g_w && 1 < a.C ? a.g_w + b * a.w_bs + 1 * a.K + k : nullptr,
      a.g_w && 2 < a.C ? a.g_w + b * a.w_bs + 2 * a.K + k : nullptr,
      a.g_w && 3 < a.C ? a.g_w + b * a.w_bs + 3 * a.K + k : nullptr};
  block_atomic_sum<2 + LX_MAX_C>(acc, dst);
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
  float* const dst[1] = {a.g_dt ? a.g_dt + b * a.dt_bs : nullptr};
  block_atomic_sum<1>(acc, dst);
}

// lambda[b][cin k] += correlation of dU_k with kernel k; dspec[b][k] += rfft2(dU_k) conj(FA[b][cin k])
template <int H, int W>
int adjoint_potential_2d(const lx_rule* rule, int64_t B, const float2* Kh, int64_t k_bs, const float* dU,
                         const float2* FA, float2* dspec, float* lam, float2* work, hipStream_t st) {
  if (int rc = tw_pair<H, W>(st)) return rc;
  constexpr int WH = W / 2 + 1;
  const int K = rule->K;
  // the K gradient images of a world are transformed as K "channels" with kernel k on image k
  launch_rows_fwd<H, W>(dU, work, B * K, st);
  ColArgs ca = col_args();
  ca.S = work; ca.U = work;  // each column is read whole before it is written back
  ca.Kh = Kh; ca.k_bstride = k_bs; ca.ncols = B * K * int64_t(WH); ca.C = K; ca.K = K; ca.conjk = 1;
  for (int k = 0; k < LX_MAX_K; ++k) {
    ca.cin[k] = k < K ? k : -1;
    ca.acc_cin[k] = k < K ? rule->cin[k] : 0;
  }
  ca.FAin = FA; ca.dspec = FA ? dspec : nullptr; ca.C_fa = rule->C;
  launch_cols<H, W>(ca, st);
  RowArgs ra = row_args(rule, B, H);
  ra.U = work; ra.A = lam; ra.A_out = lam;  // lambda_c += sum over kernels fed by c (in place, per element)
  ra.update = 1; ra.sel = 1; ra.average = 0; ra.state_fn = LX_STATE_SIMPLE;
  launch_rows_inv<H, W>(ra, st);
  return launched("adjoint potential");
}

}  // namespace

extern "C" int lx_kernel_spectrum(int32_t ndim, const int32_t* dims, int64_t n, const float* k_d, void* K_full_d,
                                  void* work_d, void* stream) {
  if (!dims || !k_d || !K_full_d || !work_d || n < 0) return lx_fail(LX_EINVAL, "bad arguments");
  if (ndim != 2) return lx_fail(LX_ENOTSUP, "lx_kernel_spectrum: 3-D kernels are not built yet");
  if (n == 0) return LX_OK;
  float2* full = static_cast<float2*>(K_full_d);
  float2* work = static_cast<float2*>(work_d);
  hipStream_t st = (hipStream_t)stream;
  return dispatch_hw(dims[0], dims[1], [&](auto h, auto w) {
    return spectrum_2d<decltype(h)::value, decltype(w)::value>(n, k_d, full, work, st);
  });
}

extern "C" int lx_rfft2(int32_t ndim, const int32_t* dims, int64_t n, const float* x_d, void* X_d, void* stream) {
  if (!dims || !x_d || !X_d || n < 0) return lx_fail(LX_EINVAL, "bad arguments");
  if (ndim != 2) return lx_fail(LX_ENOTSUP, "lx_rfft2: 3-D arrays are not built yet");
  if (n == 0) return LX_OK;
  float2* X = static_cast<float2*>(X_d);
  hipStream_t st = (hipStream_t)stream;
  return dispatch_hw(dims[0], dims[1], [&](auto h, auto w) {
    return rfft2_2d<decltype(h)::value, decltype(w)::value>(n, x_d, X, st);
  });
}

extern "C" int lx_irfft2(int32_t ndim, const int32_t* dims, int64_t n, const void* X_d, float* x_d, void* work_d,
                         void* stream) {
  if (!dims || !x_d || !X_d || !work_d || n < 0) return lx_fail(LX_EINVAL, "bad arguments");
  if (ndim != 2) return lx_fail(LX_ENOTSUP, "lx_irfft2: 3-D arrays are not built yet");
  if (n == 0) return LX_OK;
  const float2* X = static_cast<const float2*>(X_d);
  float2* work = static_cast<float2*>(work_d);
  hipStream_t st = (hipStream_t)stream;
  return dispatch_hw(dims[0], dims[1], [&](auto h, auto w) {
    return irfft2_2d<decltype(h)::value, decltype(w)::value>(n, X, x_d, work, st);
  });
}

extern "C" int lx_adjoint_pointwise(const lx_rule* rule, int64_t B, const float* state_d, const float* field_d,
                                    const float* potential_d, const float* gf_d, int64_t gf_bstride, const float* w_d,
                                    int64_t w_bstride, const float* dt_d, int64_t dt_bstride, const float* lambda_d,
                                    const float* dfield_d, const float* dpot_in_d, int32_t ste, float* lambda_out_d,
                                    float* dpot_d, float* grad_gf_d, float* grad_w_d, float* grad_dt_d,
                                    void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!state_d || !field_d || !potential_d || !gf_d || !w_d || !dt_d || !lambda_d || !lambda_out_d || !dpot_d)
    return lx_fail(LX_EINVAL, "NULL pointer");
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
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
  const unsigned per_img = unsigned(std::min<int64_t>((a.cells + 1023) / 1024, 256));
  hipStream_t st = (hipStream_t)stream;
  // growth first: it reads lambda, which the state kernel may overwrite in place
  hipLaunchKernelGGL(adj_growth_kernel, dim3(per_img, unsigned(B * rule->K)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(adj_state_kernel, dim3(per_img, unsigned(B * rule->C)), dim3(256), 0, st, a);
  return launched("adjoint pointwise");
}

extern "C" int lx_half_to_full(int32_t ndim, const int32_t* dims, int64_t n, const void* half_d, void* full_d,
                               float scale, void* stream) {
  if (!dims || !half_d || !full_d || n < 0) return lx_fail(LX_EINVAL, "bad arguments");
  if (ndim != 2) return lx_fail(LX_ENOTSUP, "lx_half_to_full: 3-D arrays are not built yet");
  if (dims[0] < 1 || dims[1] < 2) return lx_fail(LX_EINVAL, "dims %dx%d", dims[0], dims[1]);
  if (n == 0) return LX_OK;
  hipLaunchKernelGGL(expand_half_kernel, dim3(grid_for(n * dims[0] * int64_t(dims[1]), 256)), dim3(256), 0,
                     (hipStream_t)stream, static_cast<const float2*>(half_d), static_cast<float2*>(full_d), n,
                     dims[0], dims[1], scale);
  return launched("half to full");
}

extern "C" int lx_adjoint_potential(const lx_rule* rule, int64_t B, const void* K_half_d, int64_t k_bstride,
                                    const float* dpot_d, const void* state_spec_d, void* dspec_d, float* lambda_d,
                                    void* work_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!K_half_d || !dpot_d || !lambda_d || !work_d) return lx_fail(LX_EINVAL, "NULL pointer");
  if ((state_spec_d == nullptr) != (dspec_d == nullptr))
    return lx_fail(LX_EINVAL, "state_spec and dspec are given together or not at all");
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  if (rule->ndim != 2) return lx_fail(LX_ENOTSUP, "lx_adjoint_potential: 3-D worlds are not built yet");
  const float2* Kh = static_cast<const float2*>(K_half_d);
  const float2* FA = static_cast<const float2*>(state_spec_d);
  float2* dspec = static_cast<float2*>(dspec_d);
  float2* work = static_cast<float2*>(work_d);
  hipStream_t st = (hipStream_t)stream;
  return dispatch_hw(rule->dims[0], rule->dims[1], [&](auto h, auto w) {
    return adjoint_potential_2d<decltype(h)::value, decltype(w)::value>(rule, B, Kh, k_bstride, dpot_d, FA, dspec,
                                                                        lambda_d, work, st);
  });
}

extern "C" int lx_workspace_size(const lx_rule* rule, int64_t B, size_t* bytes) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!bytes || B < 0) return lx_fail(LX_EINVAL, "bad arguments");
  int64_t half = 1;
  for (int i = 0; i < rule->ndim - 1; ++i) half *= rule->dims[i];
  half *= rule->dims[rule->ndim - 1] / 2 + 1;
  *bytes = size_t(B) * size_t(rule->C + rule->K) * size_t(half) * sizeof(float2);
  return LX_OK;
}

extern "C" int lx_update(const lx_rule* rule, const float* state_d, int64_t B, const void* K_half_d,
                         int64_t k_bstride, const float* gf_d, int64_t gf_bstride, const float* w_d, int64_t w_bstride,
                         const float* dt_d, int64_t dt_bstride, float* state_out_d, float* field_d,
                         float* potential_d, void* work_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!state_d || !K_half_d || !gf_d || !w_d || !dt_d || !state_out_d || !work_d)
    return lx_fail(LX_EINVAL, "NULL pointer");
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  if (state_out_d == state_d) return lx_fail(LX_EINVAL, "lx_update: state_out must not alias state");
  return plain_step(rule, state_d, B, K_half_d, k_bstride, gf_d, gf_bstride, w_d, w_bstride, dt_d, dt_bstride,
                    state_out_d, field_d, potential_d, work_d, (hipStream_t)stream);
}

extern "C" int lx_potential(const lx_rule* rule, const float* state_d, int64_t B, const void* K_half_d,
                            int64_t k_bstride, float* potential_d, void* work_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!state_d || !K_half_d || !potential_d || !work_d) return lx_fail(LX_EINVAL, "NULL pointer");
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  // selection mode: kernel k's potential is written by the row pass of its input channel; no update
  return plain_step(rule, state_d, B, K_half_d, k_bstride, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, nullptr,
                    potential_d, work_d, (hipStream_t)stream, 1);
}

extern "C" int lx_stats(const lx_rule* rule, const float* cells_d, const float* field_d, const float* potential_d,
                        int64_t B, float R, float T, void* carry_d, float* stats_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!cells_d || !field_d || !potential_d || !carry_d || !stats_d) return lx_fail(LX_EINVAL, "NULL pointer");
  if (B < 0 || !(R > 0.f) || !(T > 0.f)) return lx_fail(LX_EINVAL, "bad B / R / T");
  if (B == 0) return LX_OK;
  StatArgs sa;
  sa.cells = cells_d; sa.field = field_d; sa.pot = potential_d;
  sa.carry = static_cast<lx_carry*>(carry_d);
  sa.stats = stats_d; sa.B = B; sa.ndim = rule->ndim;
  for (int i = 0; i < 3; ++i) sa.dims[i] = i < rule->ndim ? rule->dims[i] : 1;
  sa.C = rule->C; sa.K = rule->K;
  sa.R2 = (float)((double)R * (double)R); sa.R = R; sa.dt = (float)(1.0 / (double)T);
  hipLaunchKernelGGL(stats_kernel, dim3((unsigned)B), dim3(256), 0, (hipStream_t)stream, sa);
  return launched("stats");
}
