// Internal declarations shared by the HIP translation units of libleniax_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/leniax_amd.h"
#include "lx_device.h"

#define LX_EPS 1e-7f  // leniax/constant.py:8
#define LX_NB_CHARS 112  // leniax/constant.py:13: 25 + 25 + 62 characters of the cell codec
#define LX_MONOTONIC_STOP_STEP 128  // leniax/statistics.py:284
#define LX_MASS_VOLUME_THRESHOLD 10.f  // leniax/statistics.py:313
#define LX_MASS_VOLUME_STOP_STEP 128  // leniax/statistics.py:314

// Per-step statistic slots, in the order of leniax/statistics.py:102-115
// (channel masses last).
enum : int {
  LX_ST_MASS = 0,
  LX_ST_MASS_VOLUME,
  LX_ST_MASS_DENSITY,
  LX_ST_GROWTH,
  LX_ST_GROWTH_VOLUME,
  LX_ST_GROWTH_DENSITY,
  LX_ST_MASS_SPEED,
  LX_ST_MASS_ANGLE_SPEED,
  LX_ST_MASS_GROWTH_DIST,
  LX_ST_INERTIA,
  LX_ST_POTENTIAL_VOLUME,
  LX_ST_CHANNEL_MASS0,
};

// Per-world statistics carry (leniax/runner.py:271-292) and the online form of
// the heuristic carry of leniax/statistics.py:144-194.
struct lx_carry {
  int32_t shift[3];
  float centroid[3];
  float angle;
  float init_cm[LX_MAX_C];
  float prev_mass;
  float prev_sign;
  int32_t mono;
  int32_t vol;
  int32_t cont;
  int32_t n;
  int32_t t;
  int32_t pad;
};

// Carry record of the HBM-path statistics: the carry plus one slot of partial sums per
// block of the per-world reduction (summed in slot order: deterministic).
#define LX_STAT_PARTS 16
#define LX_NSUM (14 + LX_MAX_C)
// row-pair record of lx_step_stats: sA, sG, nA, nG, nU, sAy[3], sGy[3], sAyy[3], 2 unused
#define LX_PAIR_REC 16
struct lx_stat_carry {
  lx_carry c;
  float part[LX_STAT_PARTS][LX_NSUM];
};

// Scratch of lx_run: 4 counters (queue heads of the two launches, survivor count), the
// survivor list, one parked carry per world and the steps each world ran (for the statistics
// pass when the caller does not ask for them).
struct lx_run_scratch {
  size_t list, park, steps, total;
};
inline lx_run_scratch lx_run_scratch_layout(int64_t B) {
  lx_run_scratch l;
  l.list = 16;
  l.park = (l.list + size_t(B) * 4 + 255) / 256 * 256;
  l.steps = (l.park + size_t(B) * sizeof(lx_carry) + 255) / 256 * 256;
  l.total = l.steps + size_t(B) * 4;
  return l;
}

// Per-world reductions one step's statistics need.  Coordinates are the
// re-centred ones of leniax/statistics.py:28-33,64 (index - shift mod size - size//2).
struct lx_sums {
  float sA, sG, nA, nG, nU;
  float sAy[3], sGy[3], sAyy[3];
  float sAc[LX_MAX_C];
};

// Python-style float remainder (jnp.remainder): result has the divisor's sign.
LX_DEV float lx_fmod_floor(float x, float y) {
  float r = fmodf(x, y);
  if (r != 0.f && ((r < 0.f) != (y < 0.f))) r += y;
  return r;
}

// One step of leniax's heuristic scan (leniax/statistics.py:144-194, carry of :186-194):
// updates the counters and the cumulative continue flag; c.n counts continue steps.
LX_DEV void lx_heuristic_step(lx_carry& c, float mass, const float* cm, float mass_volume, int C) {
  // loops run to the compile-time bounds with guards, so that per-channel and per-axis arrays
  // stay in registers (no private-memory indexing) when C and ndim are run-time values
  if (c.t == 0) {
#pragma unroll
    for (int ch = 0; ch < LX_MAX_C; ++ch)
      if (ch < C) c.init_cm[ch] = cm[ch];
    c.prev_mass = mass;
    c.prev_sign = 0.f;
    c.mono = 0;
    c.vol = 0;
    c.cont = 1;
    c.n = 0;
  }
  int cond = 1;
#pragma unroll
  for (int ch = 0; ch < LX_MAX_C; ++ch) {
    if (ch >= C) break;
    cond &= (cm[ch] >= LX_EPS) ? 1 : 0;               // min_channel_mass_heuristic :224-236
    cond &= (cm[ch] <= 3.f * c.init_cm[ch]) ? 1 : 0;  // max_channel_mass_heuristic :239-251
  }
  const float dm = mass - c.prev_mass;
  const float sign = dm > 0.f ? 1.f : (dm < 0.f ? -1.f : dm * 0.f);
  c.mono = (sign == c.prev_sign) ? c.mono + 1 : 1;  // monotonic_heuristic :287-306
  cond &= (c.mono <= LX_MONOTONIC_STOP_STEP) ? 1 : 0;
  c.vol = (mass_volume > LX_MASS_VOLUME_THRESHOLD) ? c.vol + 1 : 1;  // mass_volume_heuristic :317-333
  cond &= (c.vol <= LX_MASS_VOLUME_STOP_STEP) ? 1 : 0;
  c.cont = c.cont & cond;
  c.n += c.cont;
  c.prev_mass = mass;
  c.prev_sign = sign;
  c.t += 1;
}

// Statistics of one step (leniax/statistics.py:64-115) from its sums and the carried centroid
// and angle, which it then advances (:117-124).  out (11 + C floats) may be NULL.
// R2 = float(R**2), R, dt = float(1/T) are the config's values.
LX_DEV void lx_stats_values(const lx_sums& s, float* centroid, float& angle, int ndim, int C, float R2, float R,
                            float dt, float* out) {
  const float mass = s.sA / R2;
  const float mass_volume = s.nA / R2;
  const float mass_density = mass / (mass_volume + LX_EPS);
  const float growth_ = s.sG / R2;
  const float growth_volume = s.nG / R2;
  const float growth_density = growth_ / (growth_volume + LX_EPS);
  const float potential_volume = s.nU / R2;
  float cen[3], d[3], gc[3];
  float dist2 = 0.f, gd2 = 0.f, inertia = 0.f;
  const float m00e = s.sA + LX_EPS, g00e = s.sG + LX_EPS, m002e = s.sA * s.sA + LX_EPS;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (i >= ndim) break;
    cen[i] = s.sAy[i] / m00e;
    d[i] = cen[i] - centroid[i];
    dist2 += d[i] * d[i];
    gc[i] = s.sGy[i] / g00e;
    float e = gc[i] - cen[i];
    gd2 += e * e;
    inertia += (s.sAyy[i] - cen[i] * s.sAy[i]) / m002e;
  }
  const float dist = sqrtf(dist2);
  const float mass_speed = dist / R / dt;
  const float keep = (dist / R > 0.001f) ? 1.f : 0.f;
  const float ang = (atan2f(d[1], d[0]) * 57.29577951308232f) * keep;
  const float mass_angle_speed = (lx_fmod_floor(ang - angle + 540.f, 360.f) - 180.f) / dt;
  const float mass_growth_dist = sqrtf(gd2) / R;
  if (out) {
    out[LX_ST_MASS] = mass;
    out[LX_ST_MASS_VOLUME] = mass_volume;
    out[LX_ST_MASS_DENSITY] = mass_density;
    out[LX_ST_GROWTH] = growth_;
    out[LX_ST_GROWTH_VOLUME] = growth_volume;
    out[LX_ST_GROWTH_DENSITY] = growth_density;
    out[LX_ST_MASS_SPEED] = mass_speed;
    out[LX_ST_MASS_ANGLE_SPEED] = mass_angle_speed;
    out[LX_ST_MASS_GROWTH_DIST] = mass_growth_dist;
    out[LX_ST_INERTIA] = inertia;
    out[LX_ST_POTENTIAL_VOLUME] = potential_volume;
#pragma unroll
    for (int ch = 0; ch < LX_MAX_C; ++ch)
      if (ch < C) out[LX_ST_CHANNEL_MASS0 + ch] = s.sAc[ch] / R2;
  }
  // carried centroid: the fractional part left after the truncating shift (:117-124)
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (i < ndim) centroid[i] = cen[i] - (float)(int)cen[i];
  angle = ang;
}

// What steers a run after one step: the carried shift (leniax/statistics.py:117-124, truncating
// shift, floor-mod wrap), the heuristics (:134-205, online) and the stop decision.
LX_DEV int lx_stats_control(const lx_sums& s, lx_carry& c, int ndim, const int32_t* dims, int C, float R2,
                            int stop_policy, int min_steps, float cells_total) {
  const float m00e = s.sA + LX_EPS;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (i >= ndim) break;
    const float cen = s.sAy[i] / m00e;
    int si = (int)cen;
    int v = (c.shift[i] + si) % dims[i];
    c.shift[i] = v < 0 ? v + dims[i] : v;
  }
  float cm[LX_MAX_C];
#pragma unroll
  for (int ch = 0; ch < LX_MAX_C; ++ch) cm[ch] = ch < C ? s.sAc[ch] / R2 : 0.f;
  lx_heuristic_step(c, s.sA / R2, cm, s.nA / R2, C);
  switch (stop_policy) {
    case LX_STOP_LENIAX:
      return (!c.cont && c.t >= min_steps) ? 1 : 0;
    case LX_STOP_MASS:
      return (s.sA == 0.f || s.nA > 0.9f * cells_total) ? 1 : 0;
    default:
      return 0;
  }
}

// Statistics, carry update and stop decision of one step.
LX_DEV int lx_stats_finish(const lx_sums& s, lx_carry& c, int ndim, const int32_t* dims, int C, float R2, float R,
                           float dt, int stop_policy, int min_steps, float cells_total, float* out) {
  lx_stats_values(s, c.centroid, c.angle, ndim, C, R2, R, dt, out);
  return lx_stats_control(s, c, ndim, dims, C, R2, stop_policy, min_steps, cells_total);
}

// Record of one fused 2-D one-channel step before the statistics pass: the world sums
// sA, sG, nA, nG, nU, sAy[0..1], sGy[0..1], sAyy[0..1] in the step's 12-float statistics slot.
LX_DEV lx_sums lx_sums_of_record(const float* r) {
  lx_sums s;
  s.sA = r[0]; s.sG = r[1]; s.nA = r[2]; s.nG = r[3]; s.nU = r[4];
  s.sAy[0] = r[5]; s.sAy[1] = r[6]; s.sAy[2] = 0.f;
  s.sGy[0] = r[7]; s.sGy[1] = r[8]; s.sGy[2] = 0.f;
  s.sAyy[0] = r[9]; s.sAyy[1] = r[10]; s.sAyy[2] = 0.f;
  s.sAc[0] = r[0];
  for (int ch = 1; ch < LX_MAX_C; ++ch) s.sAc[ch] = 0.f;
  return s;
}

// The world sums of one step from a world's row-pair records (lx_step_stats' inverse row pass:
// C x pairs records of LX_PAIR_REC floats, channel-major), summed by a block of kPairThreads
// threads: thread i takes records i, i + kPairThreads, ... in order, then the waves butterfly and
// thread 0 adds the wave sums in order (deterministic).  The result is complete in thread 0.
constexpr int kPairThreads = 256;
LX_DEV lx_sums lx_block_pair_sums(const float* __restrict__ rec, int C, int64_t pairs) {
  constexpr int NV = 14 + LX_MAX_C;  // the record's 14 sums, then the channel masses
  __shared__ float red[kPairThreads / 64][NV];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  float v[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = 0.f;
  const int64_t per_world = int64_t(C) * pairs;
  const float4* base = reinterpret_cast<const float4*>(rec);
  static_assert(LX_PAIR_REC == 16, "four float4 per record");
  for (int64_t j = tid; j < per_world; j += kPairThreads) {
    const float4 q0 = base[j * 4], q1 = base[j * 4 + 1], q2 = base[j * 4 + 2], q3 = base[j * 4 + 3];
    const float r[14] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, q3.x, q3.y};
#pragma unroll
    for (int i = 0; i < 14; ++i) v[i] += r[i];
    const int ch = int(j / pairs);
#pragma unroll
    for (int c2 = 0; c2 < LX_MAX_C; ++c2) v[14 + c2] += c2 == ch ? r[0] : 0.f;
  }
  wave_sum_n(v);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) red[wave][i] = v[i];
  }
  __syncthreads();
  lx_sums s;
  if (tid == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      float acc = 0.f;
      for (int wv = 0; wv < kPairThreads / 64; ++wv) acc += red[wv][i];
      v[i] = acc;
    }
  }
  s.sA = v[0]; s.sG = v[1]; s.nA = v[2]; s.nG = v[3]; s.nU = v[4];
  for (int d = 0; d < 3; ++d) { s.sAy[d] = v[5 + d]; s.sGy[d] = v[8 + d]; s.sAyy[d] = v[11 + d]; }
  for (int ch = 0; ch < LX_MAX_C; ++ch) s.sAc[ch] = ch < C ? v[14 + ch] : 0.f;
  return s;
}

// Block-wide sums of NV per-thread values (256-thread blocks), written to the block's slot of a
// partial-sum array: part[(blockIdx.y * gridDim.x + blockIdx.x) * NV + i], summed later in a fixed
// order (lx_adj_finish_wave, adj_ysum_kernel).
template <int NV>
LX_DEV void lx_block_partial(float (&v)[NV], float* part, bool accumulate = false) {
  __shared__ float red[NV][4];
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  float w[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) w[i] = v[i];
  wave_sum_n(w);
  if (ln == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) red[i][wv] = w[i];
  }
  __syncthreads();
  float* dst = part + (int64_t(blockIdx.y) * gridDim.x + blockIdx.x) * NV;
#pragma unroll
  for (int i = 0; i < NV; ++i)
    if (threadIdx.x == i) {
      const float sum = red[i][0] + red[i][1] + red[i][2] + red[i][3];
      dst[i] = accumulate ? dst[i] + sum : sum;
    }
}

// Lanes per row transform of the row passes (lx_fft.hip): 4 up to 64 points, 8 up to 192, 16 above.
constexpr int lx_row_tpr(int W) { return W <= 64 ? 4 : (W <= 192 ? 8 : 16); }

// The gradient scalars of a one-channel one-kernel reverse step, summed by `blocks` extra workgroups
// at the end of a later kernel's grid (the reverse step's last row pass) instead of a launch of
// their own: one wave per gradient element (dm, ds per growth row; dw; ddt), whose lanes add the
// partial sums part[b][blk][4] of the worlds that feed the element, block blk by lane blk mod 64 in
// order, then a fixed-order wave sum.  Deterministic; a world's own elements do not depend on the
// batch it is in.
struct LxAdjFinish {
  const float* part;
  int64_t B;
  int per;     // partial-sum rows per world
  int blocks;  // extra workgroups (0: none)
  float* g_gf;
  int64_t gf_bs;
  float* g_w;
  int64_t w_bs;
  float* g_dt;
  int64_t dt_bs;
};

LX_DEV int64_t lx_adj_finish_outputs(const LxAdjFinish& f) {
  const int64_t ngf = f.g_gf ? (f.gf_bs ? f.B : 1) * 2 : 0;
  const int64_t nw = f.g_w ? (f.w_bs ? f.B : 1) : 0;
  const int64_t ndt = f.g_dt ? (f.dt_bs ? f.B : 1) : 0;
  return ngf + nw + ndt;
}

// element i of the gradient scalars, by the whole wave
LX_DEV void lx_adj_finish_wave(const LxAdjFinish& f, int64_t i) {
  const int ln = threadIdx.x & 63;
  const int64_t ngf = f.g_gf ? (f.gf_bs ? f.B : 1) * 2 : 0;
  const int64_t nw = f.g_w ? (f.w_bs ? f.B : 1) : 0;
  int slot;
  int64_t bt, bs;
  float* dst;
  if (i < ngf) {
    bt = i / 2; slot = int(i - bt * 2); bs = f.gf_bs; dst = f.g_gf + bt * f.gf_bs + slot;
  } else if (i < ngf + nw) {
    bt = i - ngf; slot = 2; bs = f.w_bs; dst = f.g_w + bt * f.w_bs;
  } else {
    bt = i - ngf - nw; slot = 3; bs = f.dt_bs; dst = f.g_dt + bt * f.dt_bs;
  }
  const int64_t b0 = bs ? bt : 0, b1 = bs ? bt + 1 : f.B;
  float acc = 0.f;
  for (int64_t b = b0; b < b1; ++b)
    for (int blk = ln; blk < f.per; blk += 64) acc += f.part[(b * f.per + blk) * 4 + slot];
  acc = wave_sum(acc);
  if (ln == 0) *dst += acc;
}

// One-channel one-kernel reverse steps in pieces (lx_fft.hip), for lx_adjoint_step and a run's reverse
// pass: rows (the first row pass of a step: pointwise reverse + row transform into X; bridge: after
// the later step's last row pass, in the same kernel), lines (the passes on X between), last (the
// earliest step's last row pass; finish: the gradient scalars from the partial sums, which a run
// accumulates over its steps).  lambda is updated in place; X is the step workspace.
struct LxAdj11 {
  const lx_rule* rule;
  int64_t B;
  void* stream;
  const float *gf, *w, *dt;
  int64_t gf_bs, w_bs, dt_bs;
  const void* K_half;
  int64_t k_bs;
  int32_t ste;
  float* lam;
  float *g_gf, *g_w, *g_dt;
  void* X;
  float* part;
  int bpw;  // row-pass blocks per world (partial-sum rows per world)
};
int lx_adj11_init(const lx_rule* rule, int64_t B, const float* gf_d, int64_t gf_bstride, const float* w_d,
                  int64_t w_bstride, const float* dt_d, int64_t dt_bstride, const void* K_half_d, int64_t k_bstride,
                  int32_t ste, float* lambda_d, float* grad_gf_d, float* grad_w_d, float* grad_dt_d, void* work_d,
                  void* adj_work_d, void* stream, LxAdj11* c);
int lx_adj11_rows(const LxAdj11& c, const float* U, const float* A, bool bridge, bool accumulate);
int lx_adj11_lines(const LxAdj11& c, const void* state_spec_d, void* dspec_d);
int lx_adj11_last(const LxAdj11& c, bool finish);

// The partial-sum rows [B][per][4] (dm, ds, dw, ddt) of a one-channel one-kernel reverse step inside
// an lx_adjoint_workspace_size workspace (lx_plain.hip).
float* lx_adj_work_11(const lx_rule* r, int64_t B, void* work);

// Autotuner (lx_tune.hip): knob values by launch shape.  lx_tune_lookup gives a forced or recorded
// value (1) or nothing (0); lx_tune_should_time says whether a launcher may time its candidates now
// (tuning on, stream not capturing).
int lx_tune_mode();
int lx_tune_lookup(const char* knob, const char* key, int* value);
void lx_tune_record(const char* knob, const char* key, int value);
bool lx_tune_should_time(hipStream_t st);
// While one of these lives on a thread, lx_tune_should_time is false there: the launches it covers
// use forced, recorded or default schedules and create no events.
struct LxTuneQuiet {
  LxTuneQuiet();
  ~LxTuneQuiet();
};

// Milliseconds per call of enqueue() on stream st: one untimed call, then `reps` timed ones between
// two events (the caller's stream waits for them; < 0 on a HIP error).
template <class F>
float lx_time_ms(hipStream_t st, int reps, F&& enqueue) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1.f;
  float ms = -1.f;
  if (enqueue() == 0 && hipEventRecord(e0, st) == hipSuccess) {
    int rc = 0;
    for (int i = 0; i < reps && !rc; ++i) rc = enqueue();
    if (!rc && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
        hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
      ms /= float(reps);
    else
      ms = -1.f;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return ms;
}

// Error reporting (thread-local message), lx_api.hip
int lx_fail(int code, const char* fmt, ...);
int lx_check_rule(const lx_rule* rule);

// Statistics of one step from the row-pair records of lx_step_stats' inverse row pass
// (lx_plain.hip): pairs records per (world, channel), carry [B] lx_stat_carry, stats [B][11 + C]
int lx_pair_stats_launch(const lx_rule* rule, const float* part, int64_t B, int64_t pairs, float R, float T,
                         void* carry_d, float* stats_d, hipStream_t st);

// lx_run on the multi-pass path (lx_fft.hip): any rule lx_update takes, with retirement and
// refill from a device-side list and the step loop captured as HIP graphs
int lx_passes_scratch(const lx_rule* rule, int64_t B, int32_t slots, size_t* bytes);
int lx_passes_run(const lx_rule* rule, const lx_run_args* a, hipStream_t stream);

// Fused persistent 2-D run (lx_fused2d.hip)
// LX_ENOTSUP (with the message) for rules whose spectra are real quadrants: the adjoint passes
// read complex half spectra only
inline int lx_no_quadrant(const lx_rule* rule, const char* fn) {
  return rule->k_quadrant ? lx_fail(LX_ENOTSUP, "%s: quadrant spectra (k_quadrant = 1) are not supported here", fn)
                          : LX_OK;
}
int lx_fused2d_supported(const lx_rule* rule);
int lx_fused2d_launch(const lx_rule* rule, const lx_run_args* a, hipStream_t stream);
