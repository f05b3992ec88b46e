// Fused persistent run for 2-D N x N one-channel one-kernel worlds (N = 32, 64, 128).
//
// One workgroup owns one world at a time and steps it through max_steps Lenia
// updates (leniax/core.py:13-49 with get_potential_fft, get_field, get_state)
// plus the per-step statistics (leniax/statistics.py:35-126) and stop heuristics
// (leniax/statistics.py:134-205), then takes the next world from a device-side
// queue.  The state never leaves the chip between steps.
//
// Lane groups of TPR lanes own one length-N transform (P = N / TPR points per
// lane, fft_xlds: in-register FFT, twiddles, one transpose through the group's LDS
// scratch, in-register DFTs across the group; spectra come out with lane t holding
// frequencies (J t + j) + P k2, J = P / TPR).  Thread (r, t) holds rows 2r and 2r+1
// of the state at columns TPR*m + t.  A step:
//   rows fwd  two real rows packed as one complex row -> Z (LDS)
//   cols fwd  unpack Z into the half spectrum column c = r (columns 0 and N/2
//             are real and share one transform), FFT, x kernel spectrum
//   cols inv  inverse FFT -> Y (LDS)
//   rows inv  repack rows 2r, 2r+1 from their half spectra, inverse FFT -> U;
//             growth, weighted field, clip update, statistics sums in registers.
// LDS: one complex buffer of N*N/2 (+padding) elements reused by both transposes, and
// one TPR x (P + 2) exchange scratch per lane group.
#pragma once
#include <stdio.h>

#include "lx_internal.h"

namespace lx_fused {

template <int N>
struct F2 {
  static constexpr int TPR = N >= 128 ? 8 : 4;   // lanes per transform
  static constexpr int P = N / TPR;              // complex points per lane
  static constexpr int HALF = N / 2;
  static constexpr int NT = HALF * TPR;          // one lane group per packed row / column
  static constexpr int NW = (NT + 63) / 64;      // waves
  // packed-row spectrum stride: blocks of P padded by one, plus one so that the Z stores and
  // gathers of 32 lanes (4 rows x 8 lanes) spread over the 64 LDS banks
  static constexpr int ZRS = TPR * (P + 1) + 1;
  static constexpr int YRS = HALF + 2;           // column-spectrum row stride
  static constexpr int ZSZ = HALF * ZRS;
  static constexpr int YSZ = N * YRS;
  static constexpr int BUF = ZSZ > YSZ ? ZSZ : YSZ;
  static constexpr int NRED = 12;
  static constexpr int J = P / TPR;              // frequencies k1 per lane after the exchange
  static constexpr int XRS = xch_row(P);         // exchange scratch row stride
  static constexpr int XG = TPR * XRS;           // exchange scratch per lane group
  LX_DEV static int zidx(int hp, int k) { return hp * ZRS + (k / P) * (P + 1) + (k % P); }
  LX_DEV static int ycol(int c) { return c + (c >= N / 4 ? 1 : 0); }
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
  int32_t* queue;       // this launch's queue head
  int32_t* surv_count;  // worlds parked by the probe launch
  int32_t* surv_list;
  lx_carry* park;       // parked carries, one per world
  int limit;            // steps this launch may run per world
  int resume;           // 0: probe (or whole) launch over all worlds; 1: the parked survivors
  int state_fn;
  int average;
};

// An opaque copy: stops the compiler from sharing twiddle loads between the four
// transforms of a step (which would keep them live through the whole step).
LX_DEV int fresh(int v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}

LX_DEV float maxnan(float a, float b) { return (a != a) ? a : fmaxf(a, b); }

LX_DEV float state_update(int fn, float a, float dt, float G) {
  // leniax/core.py:245-319; products and sums round separately as in the reference
  if (fn == LX_STATE_V2) return add_nc(mul_nc(a, sub_nc(1.f, dt)), mul_nc(dt, G));
  const float v = add_nc(a, mul_nc(dt, G));
  if (fn == LX_STATE_SIMPLE) return v;
  return v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
}

// Per-world constants of the field of the one kernel, G = w g(U) / w (leniax/core.py:163-242).
// The inverse transforms leave xr = U / sc in registers (sc = 1 / (2 N^2), a power of two, so
// U = xr sc exactly and fma(xr, sc, -m) is U - m rounded once): the Gaussian's
// exp(-((U - m) / s)^2 / 2) becomes one exp2 of (U - m) gk, gk = sqrt(log2(e) / 2) / s, and
// 2 g - 1 one FMA c2 e - c1.
struct FieldK {
  float sc, m, s, gk, inv_9s2, wk, inv_wden, c1, c2;
};

template <int GF>
LX_DEV float field_of(float xr, const FieldK& f) {
  if constexpr (GF == LX_GF_GAUSSIAN || GF == LX_GF_GAUSSIAN_TARGET) {
    const float z = fmaf(xr, f.sc, -f.m) * f.gk;
    const float e = __builtin_amdgcn_exp2f(-(z * z));
    return GF == LX_GF_GAUSSIAN ? fmaf(f.c2, e, -f.c1) : f.c1 * e;
  } else if constexpr (GF == LX_GF_POLY_QUAD4) {
    const float d = fmaf(xr, f.sc, -f.m);  // U - m, one rounding as in the reference
    float o = fmaxf(0.f, 1.f - (d * d) * f.inv_9s2);
    o = o * o;
    o = o * o;
    return fmaf(f.c2, o, -f.c1);
  } else {
    return mul_nc(f.wk, growth_k<GF>(f.m, f.s, xr * f.sc)) * f.inv_wden;
  }
}

template <int N, int GF, bool HIST, bool V1>
__global__ __launch_bounds__(F2<N>::NT) void fused2d_kernel(FusedParams p) {
  using G = F2<N>;
  constexpr int P = G::P, TPR = G::TPR, NT = G::NT, NW = G::NW, HALF = G::HALF;
  constexpr int KS = HALF + 1;  // half-spectrum row length
  __shared__ float2 buf[G::BUF];
  __shared__ __align__(16) float2 xch_all[(NT / TPR) * G::XG];
  __shared__ __align__(16) float2 twl[TPR * twl_row(P)];  // lane twiddle rows (fft_xlds)
  __shared__ float red[NW][G::NRED];
  __shared__ int ctrl[4];
  __shared__ lx_carry carry;  // statistics + heuristics carry of the current world (thread 0)

  const int tid = threadIdx.x;
  // lane group of a transform: TPR consecutive lanes (the exchange's 16-byte accesses of 16
  // lanes then cover the 64 LDS banks once)
  const int t0 = tid % TPR;
  const int r0 = tid / TPR;
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int J = G::J;
  float2* const xch = xch_all + r0 * G::XG;
  const int64_t cells_per_world = int64_t(N) * N;

  for (int j = tid; j < TPR * twl_row(P); j += NT) {
    const int tt = j / twl_row(P), k = j % twl_row(P);
    twl[j] = k < P ? twl_entry(N, tt, k) : make_float2(0.f, 0.f);
  }
  // the last two waves close each step (wave 0 carries the extra work of column 0): lane 0 of
  // the last wave (FIN) keeps the carried shift, thread HT the heuristics
  constexpr int FIN = NT - 64;
  constexpr int HT = NW >= 2 ? NT - 128 : 1;
  static_assert(G::NRED >= 11, "eleven world sums per wave");

  float a0[P], a1[P];
  float2 x[P];

  for (;;) {
    if (tid == 0) {
      const int i = atomicAdd(p.queue, 1);
      if (!p.resume) ctrl[0] = i;
      else ctrl[0] = i < *p.surv_count ? p.surv_list[i] : int(p.B);
    }
    __syncthreads();
    const int64_t world = ctrl[0];
    if (world >= p.B) break;

    // a resumed world continues from its parked state (in the output buffer) and carry
    const float* A0 = (p.resume ? p.cells : p.cells0) + world * cells_per_world;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      a0[m] = A0[(2 * r0) * N + TPR * m + t0];
      a1[m] = A0[(2 * r0 + 1) * N + TPR * m + t0];
    }
    const float2* kh = p.kh + world * p.k_bstride;
    FieldK fk;
    fk.sc = 0.5f / float(N * N);
    fk.m = p.gf[world * p.gf_bstride + 0];
    fk.s = p.gf[world * p.gf_bstride + 1];
    fk.wk = p.w[world * p.w_bstride];
    fk.inv_wden = 1.f / (p.average ? fk.wk : 1.f);
    fk.inv_9s2 = 1.f / (9.f * (fk.s * fk.s));
    fk.c1 = fk.wk * fk.inv_wden;
    fk.c2 = 2.f * fk.c1;
    fk.gk = (1.f / fk.s) * 0.84932180028801904272f;  // sqrt(log2(e) / 2) / s
    const float xr_eps = LX_EPS / fk.sc;  // U > eps  <=>  xr > eps / sc (exact scaling)
    const float dtw = p.dt[world * p.dt_bstride];
    if (tid == FIN) {
      if (p.resume) {
        carry = p.park[world];
      } else {
        for (int i = 0; i < 3; ++i) { carry.shift[i] = 0; carry.centroid[i] = 0.f; }
        carry.angle = 0.f;
        carry.t = 0;
        carry.cont = 1;
        carry.n = 0;
      }
      ctrl[1] = 0;
      ctrl[2] = carry.shift[0];
      ctrl[3] = carry.shift[1];
    }
    __syncthreads();

    // Closing of step ts from the world sums its waves left in red, run by three threads on two
    // waves while the others start the next step: lanes 0-10 of the last wave store the step's
    // record (the statistics pass after the run turns records into leniax's statistics), lane 0
    // of the last wave (FIN) updates the carried shift and the mass stop, thread HT runs the
    // heuristics and the leniax stop.
    auto close_step = [&](int ts) {
      if (wave == NW - 1) {
        if (p.stats && lane < 11) {  // world sum `lane`, waves added in order
          float acc = 0.f;
#pragma unroll
          for (int wv = 0; wv < NW; ++wv) acc += red[wv][lane];
          p.stats[(world * p.max_steps + ts) * (LX_NSTAT_BASE + 1) + lane] = acc;
        }
        if (lane == 0) {
          // the carried shift (leniax/statistics.py:117-124) and the mass stop
          float sA = 0.f, nA = 0.f, sAy[2] = {0.f, 0.f};
#pragma unroll
          for (int wv = 0; wv < NW; ++wv) {
            sA += red[wv][0];
            nA += red[wv][2];
            sAy[0] += red[wv][5];
            sAy[1] += red[wv][6];
          }
          const float m00e = sA + LX_EPS;
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            // |sAy| < m00e / 2 truncates the centroid sAy / m00e to 0: no division needed
            if (!(fabsf(sAy[i]) < 0.5f * m00e)) {
              const int si = (int)(sAy[i] / m00e);
              const int v = (carry.shift[i] + si) % N;
              carry.shift[i] = v < 0 ? v + N : v;
            }
          }
          if (p.stop_policy == LX_STOP_MASS) ctrl[1] = (sA == 0.f || nA > 0.9f * float(N * N)) ? 1 : 0;
          ctrl[2] = carry.shift[0];
          ctrl[3] = carry.shift[1];
        }
      }
      if (tid == HT) {
        // heuristics (leniax/statistics.py:134-205), on another wave: counters, the cumulative
        // continue flag, N, the step count, and the leniax stop
        float sA = 0.f, nA = 0.f;
#pragma unroll
        for (int wv = 0; wv < NW; ++wv) {
          sA += red[wv][0];
          nA += red[wv][2];
        }
        const float mass = sA / p.R2;
        lx_heuristic_step(carry, mass, &mass, nA / p.R2, 1);
        if (p.stop_policy == LX_STOP_LENIAX) ctrl[1] = (!carry.cont && carry.t >= p.min_steps) ? 1 : 0;
      }
    };

    int t = carry.t;
    const int t_first = t;  // first step of this residency: nothing to close before it
    const int t_end = p.limit < p.max_steps ? p.limit : p.max_steps;
    bool stopped = false;
    while (t < t_end) {
      // The lane indices are laundered through an empty asm each step so that the
      // per-lane LDS/global addresses of the transposes are recomputed inside the
      // loop instead of being hoisted out of it and pinned in registers.
      int r = r0, q = t0;
      __asm__ volatile("" : "+v"(r), "+v"(q));

      // ---- rows forward: packed row r = A[2r] + i A[2r+1]
#pragma unroll
      for (int m = 0; m < P; ++m) x[m] = make_float2(a0[m], a1[m]);
      fft_xlds<P, TPR, false>(x, q, twl + fresh(q) * twl_row(P), xch);  // x[j TPR + k2] = Z_r[(J q + j) + P k2]
#pragma unroll
      for (int j = 0; j < J; ++j)
#pragma unroll
        for (int k2 = 0; k2 < TPR; ++k2) buf[r * G::ZRS + k2 * (P + 1) + J * q + j] = x[j * TPR + k2];
      __syncthreads();
      // the previous step is closed here, while the other waves run this step's column
      // transforms; the Y barrier below publishes its stop decision and shifts
      if (t > t_first) close_step(t - 1);

      // ---- columns forward: c = r, rows h = TPR m + q; unpack the packed rows, doubled:
      //      row 2h' = za + conj zb, row 2h'+1 = (za - conj zb)/i (the factor 2 stays with the
      //      data to the end: sc = 1 / (2 N^2)); column 0 packs the real columns 0 and N/2:
      //      (Re/Im za, Re/Im zb), undoubled (its spectrum multiply doubles it)
      {
        const int c = r;
        const bool partb = (q & 1) != 0;
        const float sig = partb ? -1.f : 1.f;
        const int ia = G::zidx(q >> 1, c);
        const int ib = G::zidx(q >> 1, c == 0 ? HALF : N - c);
        if (wave != 0) {
#pragma unroll
          for (int m = 0; m < P; ++m) {
            const float2 za = buf[ia + (TPR / 2) * m * G::ZRS];
            const float2 zb = buf[ib + (TPR / 2) * m * G::ZRS];
            const float u = fmaf(sig, zb.x, za.x), v = fmaf(-sig, zb.y, za.y);
            x[m] = make_float2(partb ? v : u, partb ? -u : v);
            if ((m & 7) == 7) __builtin_amdgcn_sched_barrier(0);
          }
        } else {
          const bool c0 = (c == 0);
#pragma unroll
          for (int m = 0; m < P; ++m) {
            const float2 za = buf[ia + (TPR / 2) * m * G::ZRS];
            const float2 zb = buf[ib + (TPR / 2) * m * G::ZRS];
            const float u = fmaf(sig, zb.x, za.x), v = fmaf(-sig, zb.y, za.y);
            const float2 d = make_float2(partb ? v : u, partb ? -u : v);
            x[m] = c0 ? make_float2(partb ? za.y : za.x, partb ? zb.y : zb.x) : d;
            if ((m & 7) == 7) __builtin_amdgcn_sched_barrier(0);
          }
        }
      }
      fft_xlds<P, TPR, false>(x, q, twl + fresh(q) * twl_row(P), xch);  // x[j TPR + k2] = F[kh = (J q + j) + P k2][c]

      // ---- multiply by the half spectrum; column 0 carries columns 0 and N/2:
      //      Q[f] = X[f] a[f] + conj(X[-f]) b[f], a = K0 + KN, b = K0 - KN (doubled, as the
      //      other columns).
      //      X[-f] lives in another lane of the group: column 0 goes through its scratch.
      {
        const float2* kcol = kh + r;
        if (wave == 0 && r == 0) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
#pragma unroll
          for (int j = 0; j < J; ++j)
#pragma unroll
            for (int k2 = 0; k2 < TPR; ++k2) xch[J * q + j + P * k2] = x[j * TPR + k2];
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
          for (int j = 0; j < J; ++j) {
#pragma unroll
            for (int k2 = 0; k2 < TPR; ++k2) {
              const int f = J * q + j + P * k2;
              const float2 pm = xch[(N - f) & (N - 1)];
              const float2 kv = kh[f * KS], kn = kh[f * KS + HALF];
              const float2 a = cadd(kv, kn), b = csub(kv, kn);
              x[j * TPR + k2] = cadd(cmul(x[j * TPR + k2], a), cmul(cconj(pm), b));
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
#pragma unroll
          for (int j = 0; j < J; ++j) {
#pragma unroll
            for (int k2 = 0; k2 < TPR; ++k2) x[j * TPR + k2] = cmul(x[j * TPR + k2], kcol[(J * q + j + P * k2) * KS]);
            if ((j & 1) == 1) __builtin_amdgcn_sched_barrier(0);
          }
        }
      }

      // ---- columns inverse -> Y
      fft_xlds<P, TPR, true>(x, q, twl + fresh(q) * twl_row(P), xch);  // x[m] = column c at row h = TPR m + q (x N)
      __syncthreads();  // every column gather of Z is done before Y overwrites it
#pragma unroll
      for (int m = 0; m < P; ++m) buf[(TPR * m + q) * G::YRS + G::ycol(r)] = x[m];
      __syncthreads();
      // the previous step ended the world: this one is not applied (the state is unchanged)
      if (ctrl[1]) {
        stopped = true;
        break;
      }
      const int sh0 = ctrl[2], sh1 = ctrl[3];

      // ---- rows inverse: repack rows 2r, 2r+1 from their half spectra, k = (J q + j) + P k2:
      //      k < N/2 read directly, k > N/2 through conj(Y[N-k]); k = 0 and N/2 are real
      {
        const int rowa = (2 * r) * G::YRS;
#pragma unroll
        for (int j = 0; j < J; ++j) {
#pragma unroll
          for (int k2 = 0; k2 < TPR; ++k2) {
            const bool mirror = k2 >= TPR / 2;
            const float sg = mirror ? -1.f : 1.f;
            const int k = J * q + j + P * k2;
            int col = mirror ? N - k : k;
            const bool edge = (j == 0) && (k2 == 0 || k2 == TPR / 2) && q == 0;  // k = 0 or N/2
            if (j == 0 && k2 == TPR / 2) col = q == 0 ? 0 : col;
            const int ia = rowa + G::ycol(col);
            const float2 ya = buf[ia];
            const float2 yb = buf[ia + G::YRS];
            float2 z = make_float2(fmaf(-sg, yb.y, ya.x), fmaf(sg, ya.y, yb.x));
            if (j == 0 && (k2 == 0 || k2 == TPR / 2)) {
              const float2 zr = k2 == 0 ? make_float2(ya.x, yb.x) : make_float2(ya.y, yb.y);
              z = edge ? zr : z;
            }
            x[j * TPR + k2] = z;
          }
          if ((j & 1) == 1) __builtin_amdgcn_sched_barrier(0);
        }
      }
      fft_xlds<P, TPR, true>(x, q, twl + fresh(q) * twl_row(P), xch);  // x[m] = (U[2r][TPR m+q], U[2r+1][TPR m+q]) / sc

      // ---- growth, field, update, statistics sums.  A lane owns rows 2r, 2r+1 at columns
      //      TPR m + q: row moments factor into per-row sums, column moments into per-column
      //      sums of the two rows.
      float sa0 = 0.f, sa1 = 0.f, sg0 = 0.f, sg1 = 0.f;  // per-row sums of A and G+
      float sAy1 = 0.f, sGy1 = 0.f, sAyy1 = 0.f;         // column (x) moments
      uint32_t nA = 0, nG = 0, nU = 0;                    // per-lane counts
      const int64_t hist_off = (int64_t(t) * p.B + world) * cells_per_world;
#pragma unroll
      for (int m = 0; m < P; ++m) {
        const int col = TPR * m + q;
        const float xc = float(((col - sh1) & (N - 1)) - N / 2);
        float am = 0.f, gpm = 0.f;
#pragma unroll
        for (int part = 0; part < 2; ++part) {
          const int row = 2 * r + part;
          const float xr = part ? x[m].y : x[m].x;
          const float a = part ? a1[m] : a0[m];
          const float Gf = field_of<GF>(xr, fk);
          const float gp = fmaxf(Gf, 0.f);
          if (part) { sa1 += a; sg1 += gp; } else { sa0 += a; sg0 += gp; }
          am += a;
          gpm += gp;
          nA += (a > LX_EPS) ? 1u : 0u;
          nG += (Gf > LX_EPS) ? 1u : 0u;
          nU += (xr > xr_eps) ? 1u : 0u;
          if constexpr (HIST) {
            if (p.cells_hist) p.cells_hist[hist_off + row * N + col] = a;
            if (p.field_hist) p.field_hist[hist_off + row * N + col] = Gf;
            if (p.pot_hist) p.pot_hist[hist_off + row * N + col] = xr * fk.sc;
          }
          float an;
          if constexpr (V1) {  // leniax/core.py:245-272: clip(A + dt G, 0, 1)
            an = fminf(fmaxf(add_nc(a, mul_nc(dtw, Gf)), 0.f), 1.f);
          } else {
            an = state_update(p.state_fn, a, dtw, Gf);
          }
          if (part) a1[m] = an; else a0[m] = an;
        }
        sAy1 = fmaf(am, xc, sAy1);
        sGy1 = fmaf(gpm, xc, sGy1);
        sAyy1 = fmaf(am * xc, xc, sAyy1);
        if ((m & 7) == 7) __builtin_amdgcn_sched_barrier(0);
      }
      const float yc0 = float(((2 * r - sh0) & (N - 1)) - N / 2);
      const float yc1 = float(((2 * r + 1 - sh0) & (N - 1)) - N / 2);
      const float sA = sa0 + sa1, sG = sg0 + sg1;
      const float sAy0 = fmaf(sa0, yc0, sa1 * yc1);
      const float sGy0 = fmaf(sg0, yc0, sg1 * yc1);
      const float sAyy0 = fmaf(sa0 * yc0, yc0, sa1 * yc1 * yc1);
      float vals[8] = {sA, sG, sAy0, sAy1, sGy0, sGy1, sAyy0, sAyy1};
      constexpr int slot[8] = {0, 1, 5, 6, 7, 8, 9, 10};
      // per-wave counts are at most 2048: nA and nG share one 32-bit sum
      uint32_t cnt[2] = {nA | (nG << 16), nU};
      wave_sum_last_n(vals);
      wave_sum_last_n(cnt);
      if (lane == 63) {
#pragma unroll
        for (int i = 0; i < 8; ++i) red[wave][slot[i]] = vals[i];
        red[wave][2] = float(cnt[0] & 0xFFFFu);
        red[wave][3] = float(cnt[0] >> 16);
        red[wave][4] = float(cnt[1]);
      }
      ++t;  // no barrier: the next step's Z barrier orders the red writes before the closing
    }

    // close the last step run (its sums are in red), unless the loop stopped on a closed one
    __syncthreads();
    if (!stopped && t > t_first) close_step(t - 1);
    __syncthreads();
    const bool finished = ctrl[1] != 0 || t >= p.max_steps;
    float* A1 = p.cells + world * cells_per_world;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      A1[(2 * r0) * N + TPR * m + t0] = a0[m];
      A1[(2 * r0 + 1) * N + TPR * m + t0] = a1[m];
    }
    if (tid == FIN) {
      if (finished) {
        if (p.n_out) p.n_out[world] = carry.n;
        if (p.steps_out) p.steps_out[world] = t;
      } else {  // still running at the probe's end: park it for the resume launch
        p.park[world] = carry;
        p.surv_list[atomicAdd(p.surv_count, 1)] = int(world);
      }
    }
    __syncthreads();
  }
}

template <int UNUSED>
__global__ void zero_i32(int32_t* p) {
  if (threadIdx.x < 4) p[threadIdx.x] = 0;
}

// The statistics of a fused run (leniax/statistics.py:64-126), from the world sums each step
// left in its record, written over that record.  One wave per world; lane k walks the steps
// [k L, (k + 1) L) in order.  A lane's starting carry (centroid and angle) follows from the two
// steps before its range, which it replays first: every lane reads those records before any
// lane of the wave writes one (the writes depend on the loads' values).
template <int UNUSED>
__global__ __launch_bounds__(64) void stats_pass_kernel(float* stats, const int32_t* steps, int64_t B, int max_steps,
                                                        float R2, float R, float dt) {
  const int64_t b = blockIdx.x;
  if (b >= B) return;
  constexpr int NS = LX_NSTAT_BASE + 1;
  const int n = steps[b] < max_steps ? steps[b] : max_steps;
  const int L = (max_steps + 63) / 64;
  const int t0 = int(threadIdx.x) * L;
  const int t1 = t0 + L < n ? t0 + L : n;
  float* rec = stats + b * int64_t(max_steps) * NS;
  float centroid[3] = {0.f, 0.f, 0.f}, angle = 0.f;
  if (t0 < n) {
    for (int t = t0 >= 2 ? t0 - 2 : 0; t < t0; ++t) {
      const lx_sums s = lx_sums_of_record(rec + int64_t(t) * NS);
      lx_stats_values(s, centroid, angle, 2, 1, R2, R, dt, nullptr);
    }
  }
  for (int t = t0; t < t1; ++t) {
    float* r = rec + int64_t(t) * NS;
    const lx_sums s = lx_sums_of_record(r);
    lx_stats_values(s, centroid, angle, 2, 1, R2, R, dt, r);
  }
}

template <int N, int GF, bool HIST, bool V1>
int launch_k(const FusedParams& p, int64_t B, hipStream_t stream) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return lx_fail(LX_EDEVICE, "hipGetDevice failed");
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return lx_fail(LX_EDEVICE, "hipDeviceGetAttribute failed");
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fused2d_kernel<N, GF, HIST, V1>, F2<N>::NT, 0) !=
          hipSuccess ||
      per_cu < 1)
    return lx_fail(LX_EDEVICE, "fused kernel does not fit on a compute unit");
  int64_t grid = int64_t(cus) * per_cu;
  if (grid > B) grid = B;
  if (grid < 1) grid = 1;
  // counters: [0] probe queue head, [1] resume queue head, [2] survivor count
  int32_t* ctr = p.queue;
  // div > 0: a probe launch of max_steps / div steps per world, then the survivors resume together
  auto launch_with = [&](int div) -> int {
    const bool probe = div > 0;
    hipLaunchKernelGGL(zero_i32<0>, dim3(1), dim3(64), 0, stream, ctr);
    FusedParams p1 = p;
    p1.queue = ctr;
    p1.surv_count = ctr + 2;
    p1.resume = 0;
    p1.limit = probe ? p.max_steps / div : p.max_steps;
    hipLaunchKernelGGL((fused2d_kernel<N, GF, HIST, V1>), dim3((unsigned)grid), dim3(F2<N>::NT), 0, stream, p1);
    if (probe) {
      FusedParams p2 = p1;
      p2.queue = ctr + 1;
      p2.resume = 1;
      p2.limit = p.max_steps;
      hipLaunchKernelGGL((fused2d_kernel<N, GF, HIST, V1>), dim3((unsigned)grid), dim3(F2<N>::NT), 0, stream, p2);
    }
    if (p.stats && B > 0)
      hipLaunchKernelGGL(stats_pass_kernel<0>, dim3((unsigned)B), dim3(64), 0, stream, p.stats, p.steps_out, B,
                         p.max_steps, p.R2, p.R, p.dt_stats);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LX_OK : lx_fail(LX_EDEVICE, "fused launch failed: %s", hipGetErrorString(e));
  };
  // the probe (worlds per CU per residency): worth it when worlds retire at uneven times and the run
  // is long.  Fixed rule: max_steps / 8.  Tuned: the fastest of no probe and / 4, / 8, / 16, timed on
  // the call's own inputs when they are not overwritten by it.
  if (p.stop_policy == LX_STOP_NONE || p.max_steps <= 256) return launch_with(0);
  char key[128];
  snprintf(key, sizeof(key), "N=%d:B=%lld:steps=%d:stop=%d:min=%d", N, (long long)B, p.max_steps, p.stop_policy,
           p.min_steps);
  int div = 8;
  if (!lx_tune_lookup("fused_probe", key, &div)) {
    div = 8;
    if (p.cells != p.cells0 && lx_tune_should_time(stream)) {
      float best = -1.f;
      for (int c : {0, 4, 8, 16}) {
        const float ms = lx_time_ms(stream, 2, [&] { return launch_with(c); });
        if (ms >= 0.f && (best < 0.f || ms < best)) {
          best = ms;
          div = c;
        }
      }
      if (best >= 0.f) lx_tune_record("fused_probe", key, div);
    }
  }
  return launch_with(div);
}

template <int N, bool HIST, bool V1>
int launch_gf(const FusedParams& p, int64_t B, int gf, hipStream_t stream) {
  switch (gf) {
    case LX_GF_POLY_QUAD4: return launch_k<N, LX_GF_POLY_QUAD4, HIST, V1>(p, B, stream);
    case LX_GF_GAUSSIAN: return launch_k<N, LX_GF_GAUSSIAN, HIST, V1>(p, B, stream);
    case LX_GF_GAUSSIAN_TARGET: return launch_k<N, LX_GF_GAUSSIAN_TARGET, HIST, V1>(p, B, stream);
    case LX_GF_STEP: return launch_k<N, LX_GF_STEP, HIST, V1>(p, B, stream);
    case LX_GF_STAIRCASE: return launch_k<N, LX_GF_STAIRCASE, HIST, V1>(p, B, stream);
    case LX_GF_TRIANGLE: return launch_k<N, LX_GF_TRIANGLE, HIST, V1>(p, B, stream);
    default: return launch_k<N, LX_GF_IDENTITY, HIST, V1>(p, B, stream);
  }
}

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
  const lx_run_scratch sl = lx_run_scratch_layout(a->B);
  char* scr = static_cast<char*>(a->scratch_d);
  p.queue = reinterpret_cast<int32_t*>(scr);
  p.surv_count = p.queue + 2;
  p.surv_list = reinterpret_cast<int32_t*>(scr + sl.list);
  p.park = reinterpret_cast<lx_carry*>(scr + sl.park);
  if (!p.steps_out) p.steps_out = reinterpret_cast<int32_t*>(scr + sl.steps);  // the statistics pass needs them
  p.limit = a->max_steps;
  p.resume = 0;
  p.state_fn = rule->state_fn;
  p.average = rule->average;
  const bool hist = a->cells_hist_d || a->field_hist_d || a->potential_hist_d;
  const int gf = rule->gf[0];
  if (rule->state_fn == LX_STATE_V1)
    return hist ? launch_gf<N, true, true>(p, a->B, gf, stream) : launch_gf<N, false, true>(p, a->B, gf, stream);
  return hist ? launch_gf<N, true, false>(p, a->B, gf, stream) : launch_gf<N, false, false>(p, a->B, gf, stream);
}

}  // namespace lx_fused
