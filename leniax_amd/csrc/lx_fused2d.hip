// Fused persistent run for 2-D N x N one-channel one-kernel worlds (N = 32, 64, 128).
//
// One workgroup of 2N threads owns one world at a time and steps it through
// max_steps Lenia updates (leniax/core.py:13-49 with get_potential_fft,
// get_field, get_state) plus the per-step statistics (leniax/statistics.py:35-126)
// and stop heuristics (leniax/statistics.py:134-205), then takes the next world
// from a device-side queue.  The state never leaves the chip between steps:
//
//   registers: thread (r, q) holds rows 2r and 2r+1 at columns 4m+q, m < N/4
//   step:  rows   -- two real rows packed as one complex row, length-N FFT
//                    over a DPP quad (fft_dist), result to LDS
//          cols   -- unpack the packed rows, length-N column FFTs over a quad
//                    (columns 0 and N/2 are real and share one transform),
//                    multiply by the kernel's half spectrum, inverse FFT, to LDS
//          rows   -- repack two rows' half spectra into one complex row,
//                    inverse FFT; growth, weighted field, clip update and the
//                    statistics sums all on the registers
//
// LDS: one complex buffer of N*N/2 (+padding) elements, reused for both transposes.
#include "lx_internal.h"

namespace {

template <int N>
struct F2 {
  static constexpr int P = N / 4;            // complex points per lane
  static constexpr int HALF = N / 2;
  static constexpr int NT = (N / 2) * 4;     // threads: one quad per packed row / column
  static constexpr int NW = (NT + 63) / 64;  // waves
  static constexpr int ZRS = 4 * (P + 1);    // packed-row spectrum stride (blocks of P padded by one)
  static constexpr int YRS = HALF + 2;       // column-spectrum row stride
  static constexpr int ZSZ = HALF * ZRS;
  static constexpr int YSZ = N * YRS;
  static constexpr int BUF = ZSZ > YSZ ? ZSZ : YSZ;
  static constexpr int NRED = 12;
  LX_DEV static int zidx(int hp, int k) { return hp * ZRS + (k / P) * (P + 1) + (k % P); }
  LX_DEV static int yidx(int h, int c) { return h * YRS + c + (c >= N / 4 ? 1 : 0); }
};

struct FusedParams {
  const float* cells0;
  float* cells;
  const float2* kh;
  int64_t k_bstride;
  const float* gf;
  int64_t gf_bstride;
  const float* w;
  int64_t w_bstride;
  const float* dt;
  int64_t dt_bstride;
  int64_t B;
  int max_steps;
  int stop_policy;
  int min_steps;
  float R2, R, dt_stats;
  float* stats;
  int32_t* n_out;
  int32_t* steps_out;
  float* cells_hist;
  float* field_hist;
  float* pot_hist;
  int32_t* queue;
  int gf_kind;
  int state_fn;
  int average;
};

// An opaque copy: stops the compiler from sharing twiddle loads between the four
// transforms of a step (which would keep 2P twiddle registers live throughout).
LX_DEV int fresh(int v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}

LX_DEV float maxnan(float a, float b) { return (a != a) ? a : fmaxf(a, b); }

LX_DEV float state_update(int fn, float a, float dt, float G) {
  // leniax/core.py:245-319; products and sums kept unfused as in the reference
  if (fn == LX_STATE_V2) return __fadd_rn(__fmul_rn(a, __fsub_rn(1.f, dt)), __fmul_rn(dt, G));
  float v = __fadd_rn(a, __fmul_rn(dt, G));
  if (fn == LX_STATE_SIMPLE) return v;
  return v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
}

template <int N>
__global__ __launch_bounds__(F2<N>::NT, 2) void fused2d_kernel(FusedParams p) {
  using G = F2<N>;
  constexpr int P = G::P, NT = G::NT, NW = G::NW, HALF = G::HALF;
  constexpr int KS = HALF + 1;  // half-spectrum row length
  __shared__ float2 buf[G::BUF];
  __shared__ float2 tw[N];
  __shared__ float red[NW][G::NRED];
  __shared__ int ctrl[4];

  const int tid = threadIdx.x;
  const int q0 = tid & 3;
  const int r0 = tid >> 2;
  const int sq0 = ((q0 & 1) << 1) | (q0 >> 1);
  const int lane = tid & 63, wave = tid >> 6;
  const float inv_n2 = 1.f / float(N * N);
  const int64_t cells_per_world = int64_t(N) * N;

  for (int j = tid; j < N; j += NT) {
    double s, c;
    sincospi(2.0 * j / N, &s, &c);
    tw[j] = make_float2((float)c, (float)-s);
  }

  float a0[P], a1[P];
  float2 x[P];
  lx_carry carry;
  const int32_t dims[3] = {N, N, 1};

  for (;;) {
    if (tid == 0) ctrl[0] = atomicAdd(p.queue, 1);
    __syncthreads();
    const int64_t world = ctrl[0];
    if (world >= p.B) break;

    const float* A0 = p.cells0 + world * cells_per_world;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      a0[m] = A0[(2 * r0) * N + 4 * m + q0];
      a1[m] = A0[(2 * r0 + 1) * N + 4 * m + q0];
    }
    const float2* kh0 = p.kh + world * p.k_bstride;
    const float gm = p.gf[world * p.gf_bstride + 0];
    const float gs = p.gf[world * p.gf_bstride + 1];
    const float wk = p.w[world * p.w_bstride];
    const float wden = p.average ? wk : 1.f;
    const float dtw = p.dt[world * p.dt_bstride];
    if (tid == 0) {
      for (int i = 0; i < 3; ++i) { carry.shift[i] = 0; carry.centroid[i] = 0.f; }
      carry.angle = 0.f;
      carry.t = 0;
      carry.cont = 1;
      carry.n = 0;
      ctrl[2] = 0;
      ctrl[3] = 0;
    }
    __syncthreads();

    int t = 0;
    while (t < p.max_steps) {
      // The lane indices are laundered through an empty asm each step so that the
      // ~200 per-lane LDS/global addresses of the transposes are recomputed inside
      // the loop instead of being hoisted out of it and pinned in VGPRs.
      int r = r0, q = q0, sq = sq0;
      const float2* kh = kh0;
      __asm__ volatile("" : "+v"(r), "+v"(q), "+v"(sq));
      const int sh0 = ctrl[2], sh1 = ctrl[3];
      // ---- forward rows: packed row r = A[2r] + i A[2r+1]
#pragma unroll
      for (int m = 0; m < P; ++m) x[m] = make_float2(a0[m], a1[m]);
      fft_dist<P, false>(x, fresh(q), tw);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) buf[r * G::ZRS + sq * (P + 1) + k1] = x[k1];
      __syncthreads();
      // ---- columns: c = r, rows h = 4m + q; unpack the packed rows
      {
        const int c = r;
        const bool c0 = (c == 0);
        const bool partb = (q & 1) != 0;
        const int ia = G::zidx(q >> 1, c);
        const int ib = G::zidx(q >> 1, c0 ? HALF : N - c);
#pragma unroll
        for (int m = 0; m < P; ++m) {
          const float2 za = buf[ia + 2 * m * G::ZRS];
          const float2 zb = buf[ib + 2 * m * G::ZRS];
          // part a (row 2h'):   ((za + conj zb)/2) ; part b (row 2h'+1): ((za - conj zb)/2i)
          // column 0 packs the real columns 0 and N/2: (Re/Im za, Re/Im zb)
          const float u1 = partb ? za.y : za.x, u2 = partb ? zb.y : zb.x;
          const float w1 = partb ? zb.x : za.y, w2 = partb ? za.x : zb.y;
          x[m] = make_float2(c0 ? u1 : 0.5f * (u1 + u2), c0 ? u2 : 0.5f * (w1 - w2));
          if ((m & 7) == 7) __builtin_amdgcn_sched_barrier(0);
        }
      }
      fft_dist<P, false>(x, fresh(q), tw);  // x[k1] = F[kh = k1 + P*sq][c]
      // ---- multiply by the half spectrum; column 0 carries columns 0 and N/2
      {
        const float2* kcol = kh + r;
        if (wave != 0) {
#pragma unroll
          for (int k1 = 0; k1 < P; ++k1) {
            x[k1] = cmul(x[k1], kcol[(k1 + P * sq) * KS]);
            if ((k1 & 7) == 7) __builtin_amdgcn_sched_barrier(0);
          }
        } else {
          const bool c0 = (r == 0);
          // Q[f] = X[f] a[f] + conj(X[-f]) b[f], a = (K0 + KN)/2, b = (K0 - KN)/2 on column 0
          {
            const float2 pm = dpp2<LX_QP(0, 1, 3, 2)>(x[0]);
            const int row = P * sq;
            const float2 kv = kcol[row * KS];
            const float2 kn = kh[row * KS + HALF];
            if (c0) {
              const float2 a = cscale(cadd(kv, kn), 0.5f), b = cscale(csub(kv, kn), 0.5f);
              x[0] = cadd(cmul(x[0], a), cmul(cconj(pm), b));
            } else {
              x[0] = cmul(x[0], kv);
            }
          }
#pragma unroll
          for (int k1 = 1; k1 <= P / 2; ++k1) {
            const int k2 = P - k1;
            const float2 p1 = dpp2<LX_QP(3, 2, 1, 0)>(x[k2]);
            const float2 p2 = dpp2<LX_QP(3, 2, 1, 0)>(x[k1]);
            const int row1 = k1 + P * sq, row2 = k2 + P * sq;
            const float2 kv1 = kcol[row1 * KS], kn1 = kh[row1 * KS + HALF];
            const float2 kv2 = kcol[row2 * KS], kn2 = kh[row2 * KS + HALF];
            float2 r1, r2;
            if (c0) {
              r1 = cadd(cmul(x[k1], cscale(cadd(kv1, kn1), 0.5f)), cmul(cconj(p1), cscale(csub(kv1, kn1), 0.5f)));
              r2 = cadd(cmul(x[k2], cscale(cadd(kv2, kn2), 0.5f)), cmul(cconj(p2), cscale(csub(kv2, kn2), 0.5f)));
            } else {
              r1 = cmul(x[k1], kv1);
              r2 = cmul(x[k2], kv2);
            }
            x[k1] = r1;
            if (k2 != k1) x[k2] = r2;
          }
        }
      }
      fft_dist<P, true>(x, fresh(q), tw);  // x[m] = column c at row h = 4m + q (x N)
      __syncthreads();
#pragma unroll
      for (int m = 0; m < P; ++m) buf[G::yidx(4 * m + q, r)] = x[m];
      __syncthreads();
      // ---- rows: repack rows 2r, 2r+1 from their half spectra, k = k1 + P*sq:
      // k <= N/2 read directly, k > N/2 through conj(Y[N-k]); k = 0 and N/2 are real.
      {
        const bool mirror = sq >= 2;
        const float sg = mirror ? -1.f : 1.f;
        const int dir = mirror ? -1 : 1;
        const int kbase = mirror ? N - P * sq : P * sq;  // column of k1 = 0 before special-casing
        const int rowa = G::yidx(2 * r, 0);
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) {
          int col = kbase + dir * k1;
          if (k1 == 0) col = (sq == 0 || sq == 2) ? 0 : P;
          const int ia = rowa + col + (col >= N / 4 ? 1 : 0);
          const float2 ya = buf[ia];
          const float2 yb = buf[ia + G::YRS];
          float2 z = make_float2(fmaf(-sg, yb.y, ya.x), fmaf(sg, ya.y, yb.x));
          if (k1 == 0) {
            const bool s0 = sq == 0, s2 = sq == 2;
            z.x = s0 ? ya.x : (s2 ? ya.y : z.x);
            z.y = s0 ? yb.x : (s2 ? yb.y : z.y);
          }
          x[k1] = z;
          if ((k1 & 7) == 7) __builtin_amdgcn_sched_barrier(0);
        }
      }
      fft_dist<P, true>(x, fresh(q), tw);  // x[m] = (U[2r][4m+q], U[2r+1][4m+q]) * N^2
      // ---- growth, field, update, statistics sums
      float sA = 0.f, sG = 0.f, nA = 0.f, nG = 0.f, nU = 0.f;
      float sAy0 = 0.f, sAy1 = 0.f, sGy0 = 0.f, sGy1 = 0.f, sAyy0 = 0.f, sAyy1 = 0.f;
      const int64_t hist_off = (int64_t(t) * p.B + world) * cells_per_world;
#pragma unroll
      for (int m = 0; m < P; ++m) {
        const int col = 4 * m + q;
        const float xc = float(((col - sh1) & (N - 1)) - N / 2);
#pragma unroll
        for (int part = 0; part < 2; ++part) {
          const int row = 2 * r + part;
          const float U = (part ? x[m].y : x[m].x) * inv_n2;
          const float a = part ? a1[m] : a0[m];
          const float g = growth(p.gf_kind, gm, gs, U);
          const float Gf = __fdiv_rn(__fmul_rn(wk, g), wden);
          const float gp = maxnan(Gf, 0.f);
          const float yc = float(((row - sh0) & (N - 1)) - N / 2);
          sA += a;
          sG += gp;
          nA += (a > LX_EPS) ? 1.f : 0.f;
          nG += (gp > LX_EPS) ? 1.f : 0.f;
          nU += (U > LX_EPS) ? 1.f : 0.f;
          sAy0 = fmaf(a, yc, sAy0);
          sAy1 = fmaf(a, xc, sAy1);
          sGy0 = fmaf(gp, yc, sGy0);
          sGy1 = fmaf(gp, xc, sGy1);
          sAyy0 = fmaf(a * yc, yc, sAyy0);
          sAyy1 = fmaf(a * xc, xc, sAyy1);
          if (p.cells_hist) p.cells_hist[hist_off + row * N + col] = a;
          if (p.field_hist) p.field_hist[hist_off + row * N + col] = Gf;
          if (p.pot_hist) p.pot_hist[hist_off + row * N + col] = U;
          const float an = state_update(p.state_fn, a, dtw, Gf);
          if (part) a1[m] = an; else a0[m] = an;
        }
        if ((m & 7) == 7) __builtin_amdgcn_sched_barrier(0);
      }
      float vals[11] = {sA, sG, nA, nG, nU, sAy0, sAy1, sGy0, sGy1, sAyy0, sAyy1};
#pragma unroll
      for (int i = 0; i < 11; ++i) {
        const float v = wave_sum(vals[i]);
        if (lane == 0) red[wave][i] = v;
      }
      __syncthreads();
      if (tid == 0) {
        lx_sums s;
        float tot[11];
        for (int i = 0; i < 11; ++i) {
          float acc = 0.f;
          for (int wv = 0; wv < NW; ++wv) acc += red[wv][i];
          tot[i] = acc;
        }
        s.sA = tot[0]; s.sG = tot[1]; s.nA = tot[2]; s.nG = tot[3]; s.nU = tot[4];
        s.sAy[0] = tot[5]; s.sAy[1] = tot[6]; s.sAy[2] = 0.f;
        s.sGy[0] = tot[7]; s.sGy[1] = tot[8]; s.sGy[2] = 0.f;
        s.sAyy[0] = tot[9]; s.sAyy[1] = tot[10]; s.sAyy[2] = 0.f;
        s.sAc[0] = tot[0];
        float* out = p.stats ? p.stats + (world * p.max_steps + t) * (LX_NSTAT_BASE + 1) : nullptr;
        const int stop = lx_stats_finish(s, carry, 2, dims, 1, p.R2, p.R, p.dt_stats, p.stop_policy, p.min_steps,
                                         float(N * N), out);
        ctrl[1] = stop;
        ctrl[2] = carry.shift[0];
        ctrl[3] = carry.shift[1];
      }
      __syncthreads();
      ++t;
      if (ctrl[1]) break;
    }

    float* A1 = p.cells + world * cells_per_world;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      A1[(2 * r0) * N + 4 * m + q0] = a0[m];
      A1[(2 * r0 + 1) * N + 4 * m + q0] = a1[m];
    }
    if (tid == 0) {
      if (p.n_out) p.n_out[world] = carry.n;
      if (p.steps_out) p.steps_out[world] = t;
    }
    __syncthreads();
  }
}

__global__ void zero_i32(int32_t* p) { *p = 0; }

template <int N>
int launch(const lx_rule* rule, const lx_run_args* a, hipStream_t stream) {
  FusedParams p;
  p.cells0 = a->cells0_d;
  p.cells = a->cells_d;
  p.kh = static_cast<const float2*>(a->K_half_d);
  p.k_bstride = a->k_bstride;
  p.gf = a->gf_d;
  p.gf_bstride = a->gf_bstride;
  p.w = a->w_d;
  p.w_bstride = a->w_bstride;
  p.dt = a->dt_d;
  p.dt_bstride = a->dt_bstride;
  p.B = a->B;
  p.max_steps = a->max_steps;
  p.stop_policy = a->stop_policy;
  p.min_steps = a->min_steps;
  p.R2 = (float)((double)a->R * (double)a->R);
  p.R = a->R;
  p.dt_stats = (float)(1.0 / (double)a->T);
  p.stats = a->stats_d;
  p.n_out = a->n_d;
  p.steps_out = a->steps_d;
  p.cells_hist = a->cells_hist_d;
  p.field_hist = a->field_hist_d;
  p.pot_hist = a->potential_hist_d;
  p.queue = a->queue_d;
  p.gf_kind = rule->gf[0];
  p.state_fn = rule->state_fn;
  p.average = rule->average;

  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return lx_fail(LX_EDEVICE, "hipGetDevice failed");
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return lx_fail(LX_EDEVICE, "hipDeviceGetAttribute failed");
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fused2d_kernel<N>, F2<N>::NT, 0) != hipSuccess ||
      per_cu < 1)
    return lx_fail(LX_EDEVICE, "fused kernel does not fit on a compute unit");
  int64_t grid = int64_t(cus) * per_cu;
  if (grid > a->B) grid = a->B;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(zero_i32, dim3(1), dim3(1), 0, stream, a->queue_d);
  hipLaunchKernelGGL(fused2d_kernel<N>, dim3((unsigned)grid), dim3(F2<N>::NT), 0, stream, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return lx_fail(LX_EDEVICE, "fused launch failed: %s", hipGetErrorString(e));
  return LX_OK;
}

}  // namespace

int lx_fused2d_supported(const lx_rule* rule) {
  if (rule->ndim != 2 || rule->C != 1 || rule->K != 1) return 0;
  if (rule->dims[0] != rule->dims[1]) return 0;
  const int n = rule->dims[0];
  return n == 32 || n == 64 || n == 128;
}

int lx_fused2d_launch(const lx_rule* rule, const lx_run_args* a, hipStream_t stream) {
  switch (rule->dims[0]) {
    case 32: return launch<32>(rule, a, stream);
    case 64: return launch<64>(rule, a, stream);
    case 128: return launch<128>(rule, a, stream);
    default: return lx_fail(LX_ENOTSUP, "fused run: unsupported world size %d", rule->dims[0]);
  }
}
