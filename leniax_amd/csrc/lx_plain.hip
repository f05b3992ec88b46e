// Plain (unfused) path: the leniax step as three HBM passes, plus the elementwise
// field / state kernels, the per-world statistics and the heuristic scan.
//
//   rows_fwd:  real rows -> half spectra along W (two rows packed per complex FFT)
//   cols:      FFT along H of every half-spectrum column, multiply by each kernel's
//              half spectrum (depthwise: kernel k reads channel cin[k]), inverse FFT
//   rows_inv:  per channel c: for every kernel k feeding c, inverse FFT along W of
//              two rows at once, growth g_k, weighted sum into G_c; then the update
// Workspace: S [B][C][H][W/2+1] and U^ [B][K][H][W/2+1] complex64.
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
};

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
#pragma unroll
  for (int m = 0; m < P; ++m) x[m] = src[int64_t(TPR * m + t) * WH];
  fft_lanes<P, TPR, false>(x, t, lane, g_tw<H>);
  const int s = bitrev_n(t, TPR);
  if (a.Kh == nullptr) {  // forward transform only, in place (spectrum of real kernels)
    float2* dst = const_cast<float2*>(a.S) + img * int64_t(H) * WH + j;
    if (live) {
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) dst[int64_t(k1 + P * s) * WH] = x[k1];
    }
    return;
  }
  for (int k = 0; k < a.K; ++k) {
    if (a.cin[k] != c) continue;
    const float2* kk = a.Kh + b * a.k_bstride + int64_t(k) * H * WH + j;
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) y[k1] = cmul(x[k1], kk[int64_t(k1 + P * s) * WH]);
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
  const float* pw = a.w + b * a.w_bs;
  const float* pg = a.gf + b * a.gf_bs;
  const int s = bitrev_n(t, TPR);
  const bool mirror = s >= TPR / 2;
  const float sg = mirror ? -1.f : 1.f;
  float G0[P], G1[P];
#pragma unroll
  for (int m = 0; m < P; ++m) G0[m] = G1[m] = 0.f;
  float den = 0.f;
  float2 x[P];
  for (int k = 0; k < a.K; ++k) {
    const float wk = pw[c * a.K + k];
    den = add_nc(den, wk);
    int owner = 0;
    for (int cc = a.C - 1; cc >= 0; --cc)
      if (pw[cc * a.K + k] != 0.f) owner = cc;
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
    const float gm = pg[2 * k], gs = pg[2 * k + 1];
    float* po0 = want_pot ? a.pot + ((b * a.K + k) * int64_t(a.H) + r2) * W : nullptr;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const float U0 = x[m].x * inv_n, U1 = x[m].y * inv_n;
      if (want_pot && live) {
        po0[TPR * m + t] = U0;
        po0[W + TPR * m + t] = U1;
      }
      if (a.update) {
        G0[m] = add_nc(G0[m], mul_nc(wk, growth(a.gfk[k], gm, gs, U0)));
        G1[m] = add_nc(G1[m], mul_nc(wk, growth(a.gfk[k], gm, gs, U1)));
      }
    }
  }
  if (!a.update || !live) return;
  const float dtw = a.dt[b * a.dt_bs];
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

template <int H, int W>
int run_plain(const lx_rule* rule, const float* A, int64_t B, const float2* Kh, int64_t k_bs, const float* gf,
              int64_t gf_bs, const float* w, int64_t w_bs, const float* dt, int64_t dt_bs, float* A_out,
              float* field, float* pot, void* work, hipStream_t st) {
  int rc = ensure_tw<H>(st);
  if (!rc) rc = ensure_tw<W>(st);
  if (rc) return rc;
  constexpr int WH = W / 2 + 1;
  const int C = rule->C, K = rule->K;
  float2* S = static_cast<float2*>(work);
  float2* U = S + B * C * int64_t(H) * WH;
  const int64_t npairs = B * C * int64_t(H / 2);
  {
    const int64_t thr = npairs * Lanes<W>::TPR;
    hipLaunchKernelGGL(rows_fwd_kernel<W>, dim3(unsigned((thr + 255) / 256)), dim3(256), 0, st, A, S, npairs, H);
  }
  {
    ColArgs ca;
    ca.S = S; ca.U = U; ca.Kh = Kh; ca.k_bstride = k_bs; ca.ncols = B * C * int64_t(WH); ca.C = C; ca.K = K;
    for (int k = 0; k < LX_MAX_K; ++k) ca.cin[k] = k < K ? rule->cin[k] : -1;
    const int64_t thr = ca.ncols * Lanes<H>::TPR;
    hipLaunchKernelGGL((cols_kernel<H, W>), dim3(unsigned((thr + 255) / 256)), dim3(256), 0, st, ca);
  }
  {
    RowArgs ra;
    ra.U = U; ra.A = A; ra.A_out = A_out; ra.field = field; ra.pot = pot; ra.gf = gf; ra.gf_bs = gf_bs; ra.w = w;
    ra.w_bs = w_bs; ra.dt = dt; ra.dt_bs = dt_bs; ra.npairs = npairs; ra.H = H; ra.C = C; ra.K = K;
    for (int k = 0; k < LX_MAX_K; ++k) ra.gfk[k] = k < K ? rule->gf[k] : 0;
    ra.state_fn = rule->state_fn; ra.average = rule->average; ra.update = A_out != nullptr;
    const int64_t thr = npairs * Lanes<W>::TPR;
    hipLaunchKernelGGL(rows_inv_kernel<W>, dim3(unsigned((thr + 255) / 256)), dim3(256), 0, st, ra);
  }
  return launched("plain step");
}

template <int H>
int dispatch_w(int W, const lx_rule* rule, const float* A, int64_t B, const float2* Kh, int64_t k_bs,
               const float* gf, int64_t gf_bs, const float* w, int64_t w_bs, const float* dt, int64_t dt_bs,
               float* A_out, float* field, float* pot, void* work, hipStream_t st) {
#define LX_W(WW) \
  case WW: return run_plain<H, WW>(rule, A, B, Kh, k_bs, gf, gf_bs, w, w_bs, dt, dt_bs, A_out, field, pot, work, st);
  switch (W) {
    LX_W(16) LX_W(32) LX_W(64) LX_W(128) LX_W(256) LX_W(512)
    default: return lx_fail(LX_ENOTSUP, "width %d", W);
  }
#undef LX_W
}

int plain_step(const lx_rule* rule, const float* A, int64_t B, const void* Kh, int64_t k_bs, const float* gf,
               int64_t gf_bs, const float* w, int64_t w_bs, const float* dt, int64_t dt_bs, float* A_out,
               float* field, float* pot, void* work, hipStream_t st) {
  if (rule->ndim != 2) return lx_fail(LX_ENOTSUP, "plain step: 3-D worlds are not built yet");
  const int H = rule->dims[0], W = rule->dims[1];
  if (!plain_size_ok(H) || !plain_size_ok(W))
    return lx_fail(LX_ENOTSUP, "plain step: world %dx%d (sizes must be powers of two in [16, 512])", H, W);
  const float2* K2 = static_cast<const float2*>(Kh);
#define LX_H(HH) \
  case HH: return dispatch_w<HH>(W, rule, A, B, K2, k_bs, gf, gf_bs, w, w_bs, dt, dt_bs, A_out, field, pot, work, st);
  switch (H) {
    LX_H(16) LX_H(32) LX_H(64) LX_H(128) LX_H(256) LX_H(512)
    default: return lx_fail(LX_ENOTSUP, "height %d", H);
  }
#undef LX_H
}

// full spectrum from the half spectrum of a real array (Hermitian extension)
__global__ void expand_half_kernel(const float2* __restrict__ half, float2* __restrict__ full, int64_t n, int H, int W) {
  const int WH = W / 2 + 1;
  const int64_t total = n * H * int64_t(W);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int kw = int(i % W);
    const int kh = int((i / W) % H);
    const int64_t img = i / (int64_t(H) * W);
    const float2* hb = half + img * int64_t(H) * WH;
    full[i] = kw <= W / 2 ? hb[int64_t(kh) * WH + kw] : cconj(hb[int64_t((H - kh) % H) * WH + (W - kw)]);
  }
}

template <int H, int W>
int spectrum_2d(int64_t n, const float* k, float2* full, float2* work, hipStream_t st) {
  int rc = ensure_tw<H>(st);
  if (!rc) rc = ensure_tw<W>(st);
  if (rc) return rc;
  constexpr int WH = W / 2 + 1;
  const int64_t npairs = n * int64_t(H / 2);
  hipLaunchKernelGGL(rows_fwd_kernel<W>, dim3(unsigned((npairs * Lanes<W>::TPR + 255) / 256)), dim3(256), 0, st, k,
                     work, npairs, H);
  ColArgs ca;
  ca.S = work; ca.U = nullptr; ca.Kh = nullptr; ca.k_bstride = 0; ca.ncols = n * int64_t(WH); ca.C = 1; ca.K = 0;
  for (int i = 0; i < LX_MAX_K; ++i) ca.cin[i] = -1;
  hipLaunchKernelGGL((cols_kernel<H, W>), dim3(unsigned((ca.ncols * Lanes<H>::TPR + 255) / 256)), dim3(256), 0, st,
                     ca);
  hipLaunchKernelGGL(expand_half_kernel, dim3(grid_for(n * H * int64_t(W), 256)), dim3(256), 0, st, work, full, n, H,
                     W);
  return launched("kernel spectrum");
}

template <int H>
int spectrum_w(int W, int64_t n, const float* k, float2* full, float2* work, hipStream_t st) {
  switch (W) {
    case 16: return spectrum_2d<H, 16>(n, k, full, work, st);
    case 32: return spectrum_2d<H, 32>(n, k, full, work, st);
    case 64: return spectrum_2d<H, 64>(n, k, full, work, st);
    case 128: return spectrum_2d<H, 128>(n, k, full, work, st);
    case 256: return spectrum_2d<H, 256>(n, k, full, work, st);
    case 512: return spectrum_2d<H, 512>(n, k, full, work, st);
    default: return lx_fail(LX_ENOTSUP, "width %d", W);
  }
}

}  // namespace

extern "C" int lx_kernel_spectrum(int32_t ndim, const int32_t* dims, int64_t n, const float* k_d, void* K_full_d,
                                  void* work_d, void* stream) {
  if (!dims || !k_d || !K_full_d || !work_d || n < 0) return lx_fail(LX_EINVAL, "bad arguments");
  if (ndim != 2) return lx_fail(LX_ENOTSUP, "lx_kernel_spectrum: 3-D kernels are not built yet");
  const int H = dims[0], W = dims[1];
  if (!plain_size_ok(H) || !plain_size_ok(W)) return lx_fail(LX_ENOTSUP, "kernel spectrum: size %dx%d", H, W);
  if (n == 0) return LX_OK;
  float2* full = static_cast<float2*>(K_full_d);
  float2* work = static_cast<float2*>(work_d);
  hipStream_t st = (hipStream_t)stream;
  switch (H) {
    case 16: return spectrum_w<16>(W, n, k_d, full, work, st);
    case 32: return spectrum_w<32>(W, n, k_d, full, work, st);
    case 64: return spectrum_w<64>(W, n, k_d, full, work, st);
    case 128: return spectrum_w<128>(W, n, k_d, full, work, st);
    case 256: return spectrum_w<256>(W, n, k_d, full, work, st);
    case 512: return spectrum_w<512>(W, n, k_d, full, work, st);
    default: return lx_fail(LX_ENOTSUP, "height %d", H);
  }
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
  // potential only: every kernel is "owned" by channel 0 through unit weights
  static thread_local float* ones = nullptr;
  static thread_local int64_t ones_n = 0;
  const int64_t need = int64_t(rule->C) * rule->K + 2 * rule->K + 1;
  if (ones_n < need) {
    if (ones) (void)hipFree(ones);
    if (hipMalloc(&ones, need * sizeof(float)) != hipSuccess) return lx_fail(LX_ENOMEM, "hipMalloc failed");
    ones_n = need;
  }
  if (hipMemsetAsync(ones, 0, need * sizeof(float), (hipStream_t)stream) != hipSuccess)
    return lx_fail(LX_EDEVICE, "memset failed");
  // weights all zero -> every kernel's owner is channel 0; rows_inv writes U only (A_out == NULL)
  return plain_step(rule, state_d, B, K_half_d, k_bstride, ones, 0, ones, 0, ones, 0, nullptr, nullptr, potential_d,
                    work_d, (hipStream_t)stream);
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
