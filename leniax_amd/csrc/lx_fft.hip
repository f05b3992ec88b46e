// The leniax step on the HBM-resident path, for 2-D and 3-D worlds, as separable FFT passes
// over the Hermitian half spectrum, plus the plain transforms and the adjoint of the potential.
//
//   rows_fwd:  real rows (last axis, length W) -> half spectra along W, two rows per complex FFT
//   lines:     the FFT of every line of the half spectrum along one leading axis.  mode 0 (the
//              convolution axis): forward, x every kernel's half spectrum, inverse; mode 1:
//              forward in place; mode 2: inverse (in place or into another buffer)
//   rows_inv:  per channel c: inverse row FFT of every kernel feeding c (two rows at once),
//              growth g_k, weighted sum into G_c, then the update
//
//   2-D [H][W]:    rows_fwd, lines(H, conv), rows_inv                            (3 passes)
//   3-D [D][H][W]: rows_fwd, lines(H, fwd), lines(D, conv), lines(H, inv), rows_inv  (5 passes)
//
// Transform lengths: 2^k in [16, 512] and 3 * 2^k in [48, 384].  A length-L transform is
// spread over a group of TPR lanes of one wave (P = L / TPR points per lane): an in-register
// FFT of P points, a twiddle multiply and the cross-lane step (a transpose through LDS when
// P is a multiple of TPR, else a radix-2 network across the group through DPP).
// Workspace of a step: S [B][C][half] and U [B][K][half] complex64, half = prod(dims[:-1]) (W/2+1).
#include <stdio.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <type_traits>
#include <vector>

#include "lx_internal.h"

namespace {

// lanes per transform of the row passes: 8 from 96 points up (96 = 8 lanes x 12 points through
// the DPP network rather than 4 x 24 through LDS: half the registers of rows_inv, 3 waves per
// SIMD instead of 2)
constexpr int row_tpr(int L) { return L <= 64 ? 4 : (L <= 192 ? 8 : 16); }
// lanes per transform of the line passes: as the rows, except the plain forward and inverse
// passes (modes 1 and 2) at 96 points, which take 4 x 24 through LDS: they stay at 4 waves per
// SIMD there (101 and 112 VGPRs) and run 10% faster than on the network (the convolution pass,
// at 173 VGPRs with 4 lanes, is as fast on 8)
constexpr int line_tpr(int L, int mode) { return L == 96 && mode != 0 ? 4 : row_tpr(L); }

template <int L, int T = row_tpr(L)>
struct Lanes {
  static constexpr int TPR = T;
  static constexpr int P = L / TPR;                                // points per lane
  static constexpr bool XL = P % TPR == 0;                         // cross-lane step through LDS
  static constexpr int XRS = xch_row(P);                           // group scratch row stride
  static constexpr int XCH = 256 * XRS;                            // scratch of a 256-thread block
  static_assert(TPR == 4 || TPR == 8 || TPR == 16, "lane groups of 4, 8 or 16");
  static_assert(TPR * XRS >= L + 2, "a group's scratch holds one transform plus two");
};

template <int V>
using IntC = std::integral_constant<int, V>;

bool len_ok(int n) {
  switch (n) {
    case 16: case 32: case 48: case 64: case 96: case 128: case 192: case 256: case 384: case 512: return true;
    default: return false;
  }
}

// calls f(IntC<L>) for a supported transform length
template <typename F>
int dispatch_len(int L, F&& f) {
  switch (L) {
    case 16: return f(IntC<16>{});
    case 32: return f(IntC<32>{});
    case 48: return f(IntC<48>{});
    case 64: return f(IntC<64>{});
    case 96: return f(IntC<96>{});
    case 128: return f(IntC<128>{});
    case 192: return f(IntC<192>{});
    case 256: return f(IntC<256>{});
    case 384: return f(IntC<384>{});
    case 512: return f(IntC<512>{});
    default: return lx_fail(LX_ENOTSUP, "transform length %d (supported: 2^k in [16, 512], 3*2^k in [48, 384])", L);
  }
}

// twiddle tables in global memory, one per length and device: exp(-2 pi i j / L) for the DPP
// network, and the lane rows of fft_xlds (row t, entry k = exp(-2 pi i t k / L))
template <int L>
__device__ float2 g_tw[L];
template <int L, int T>
__device__ __align__(16) float2 g_twl[T * twl_row(L / T)];

template <int L, int T>
LX_DEV void init_twl() {
  constexpr int P = L / T;
  for (int j = threadIdx.x; j < T * twl_row(P); j += blockDim.x) {
    const int t = j / twl_row(P), k = j % twl_row(P);
    g_twl<L, T>[j] = k < P ? twl_entry(L, t, k) : make_float2(0.f, 0.f);
  }
}

template <int L>
__global__ void init_tw_kernel() {
  for (int j = threadIdx.x; j < L; j += blockDim.x) {
    double sn, cs;
    sincospi(2.0 * j / L, &sn, &cs);
    g_tw<L>[j] = make_float2((float)cs, (float)-sn);
  }
  init_twl<L, row_tpr(L)>();
  if constexpr (line_tpr(L, 1) != row_tpr(L)) init_twl<L, line_tpr(L, 1)>();
}

constexpr int kMaxDevices = 64;

// The tables of length L on the current device, built once per process and device, before the
// first kernel that reads them.  The build runs on an internal stream and is waited for, so that
// every later call, on any stream or thread, finds the tables complete.  It is refused while the
// caller's stream is being captured into a graph: a build recorded there would run only when the
// graph does, leaving the eager calls in between with empty tables.
template <int L>
int ensure_tw(hipStream_t stream) {
  static std::atomic<bool> ready[kMaxDevices];
  static std::mutex mu;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return lx_fail(LX_EDEVICE, "hipGetDevice failed or device index out of range");
  if (ready[dev].load(std::memory_order_acquire)) return LX_OK;
  std::lock_guard<std::mutex> lock(mu);
  if (ready[dev].load(std::memory_order_relaxed)) return LX_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (stream && hipStreamIsCapturing(stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
    return lx_fail(LX_EINVAL, "transform length %d is first used inside a stream capture: make one call of the "
                   "same shape outside the capture first (it builds the twiddle tables)", L);
  hipStream_t own = nullptr;
  if (hipStreamCreateWithFlags(&own, hipStreamNonBlocking) != hipSuccess)
    return lx_fail(LX_EDEVICE, "twiddle init: stream creation failed");
  hipLaunchKernelGGL(init_tw_kernel<L>, dim3(1), dim3(256), 0, own);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(own);
  hipStreamDestroy(own);
  if (e != hipSuccess) return lx_fail(LX_EDEVICE, "twiddle init: %s", hipGetErrorString(e));
  ready[dev].store(true, std::memory_order_release);
  return LX_OK;
}

int launched(const char* what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LX_OK : lx_fail(LX_EDEVICE, "%s launch: %s", what, hipGetErrorString(e));
}

unsigned blocks_for(int64_t threads) { return unsigned((threads + 255) / 256); }

// Frequency held by register e of lane t after a forward group transform: with the LDS
// exchange e = j TPR + k2 holds (J t + j) + P k2; with the DPP network e holds e + P bitrev(t).
template <int L, int T = row_tpr(L)>
LX_DEV int freq_of(int t, int e) {
  using LC = Lanes<L, T>;
  if constexpr (LC::XL) {
    constexpr int J = LC::P / LC::TPR;
    return J * t + e / LC::TPR + LC::P * (e % LC::TPR);
  } else {
    return e + LC::P * bitrev_n(t, LC::TPR);
  }
}

// Length-L transform over a lane group: natural order (data[TPR m + t] in x[m]) <-> the
// frequency layout of freq_of; xch is the group's LDS scratch.
template <int L, bool INV, int T = row_tpr(L)>
LX_DEV void fft_group(float2* x, int t, int lane, float2* xch) {
  using LC = Lanes<L, T>;
  if constexpr (LC::XL) fft_xlds<LC::P, LC::TPR, INV>(x, t, g_twl<L, T> + t * twl_row(LC::P), xch);
  else fft_lanes<LC::P, LC::TPR, INV>(x, t, lane, g_tw<L>);
}

LX_DEV void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// XCD-aware block order.  The hardware deals a grid's workgroups to the 8 XCDs in turn (block b
// runs on XCD b % 8 as that XCD's (b / 8)-th block); each XCD has its own L2.  The logical block
// of b is chosen so that every XCD runs one contiguous range of logical blocks: neighbouring
// blocks, whose lines share 128-byte cache lines (a column group's 16 columns straddle two lines
// of each 1032-byte half-spectrum row), then meet in the same L2.  A bijection of [0, nb).
LX_DEV int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t per = nb / 8, rem = nb % 8, x = b % 8, i = b / 8;
  return x * per + (x < rem ? x : rem) + i;
}

// ---------------------------------------------------------------- geometry
struct Geom {
  int ndim, W, WH;
  int64_t rows;   // rows of length W per image: prod(dims[:-1])
  int64_t half;   // complex elements of one image's half spectrum: rows * WH
  int naxes;      // leading axes: 1 (2-D) or 2 (3-D); axis 0 is the convolution axis
  int L[2], A[2];
  int64_t a_stride[2], st[2];  // line base = img * half + (a * a_stride + j), element stride st
};

int make_geom(int ndim, const int32_t* dims, Geom& g) {
  if (ndim != 2 && ndim != 3) return lx_fail(LX_EINVAL, "ndim %d", ndim);
  for (int i = 0; i < ndim; ++i)
    if (!len_ok(dims[i]))
      return lx_fail(LX_ENOTSUP, "world size %d on axis %d (supported: 2^k in [16, 512], 3*2^k in [48, 384])",
                     dims[i], i);
  g.ndim = ndim;
  g.W = dims[ndim - 1];
  g.WH = g.W / 2 + 1;
  if (ndim == 2) {
    const int H = dims[0];
    g.rows = H;
    g.naxes = 1;
    g.L[0] = H; g.A[0] = 1; g.a_stride[0] = 0; g.st[0] = g.WH;
  } else {
    const int D = dims[0], H = dims[1];
    g.rows = int64_t(D) * H;
    g.naxes = 2;
    g.L[0] = D; g.A[0] = H; g.a_stride[0] = g.WH; g.st[0] = int64_t(H) * g.WH;
    g.L[1] = H; g.A[1] = D; g.a_stride[1] = int64_t(H) * g.WH; g.st[1] = g.WH;
  }
  g.half = g.rows * g.WH;
  return LX_OK;
}

// ---------------------------------------------------------------- batched-run indirection
// The kernels of lx_run's multi-pass path (LIST = true) step the worlds of a live list: slot s of
// the list holds world list[s], slots >= *nlive are empty.  The per-step temporaries (spectra S,
// U, the row-pair records) are indexed by slot; what persists (the state, carry, parameters,
// spectra, statistics) by world.  World b's state after t of its steps lives in X0 (t even) or X1
// (t odd), t = carry[b].t: worlds advance in lockstep inside a step but start at different steps.
struct RunView {
  const int32_t* list;
  const int32_t* nlive;
  const lx_carry* carry;
  float* X0;
  float* X1;
  int64_t world_elems;  // C * cells floats per world
};

// the state buffer holding world b's current state
LX_DEV float* run_state(const RunView& v, int64_t b, bool next) {
  const bool odd = ((v.carry[b].t & 1) != 0) != next;
  return (odd ? v.X1 : v.X0) + b * v.world_elems;
}

// ---------------------------------------------------------------- rows forward
// Group = one pair of rows packed as z = a + i b; after the transform Z goes through the group's
// scratch in frequency order and comes out as the two half spectra, A[k] = (Z[k] + conj Z[W-k])/2,
// B[k] = (Z[k] - conj Z[W-k])/2i, stored with consecutive lanes on consecutive k.
// LIST: npairs is the pairs of one slot (C rows / 2); the state comes from the slot's world.
template <int W, bool LIST>
__global__ __launch_bounds__(256) void rows_fwd_kernel(const float* __restrict__ A, float2* __restrict__ S,
                                                       int64_t npairs, RunView rv) {
  using LC = Lanes<W>;
  constexpr int TPR = LC::TPR, P = LC::P, WH = W / 2 + 1;
  __shared__ __align__(16) float2 xch_all[LC::XCH];
  const int lane = threadIdx.x & 63;
  const int t = threadIdx.x % TPR;
  float2* xch = xch_all + (threadIdx.x / TPR) * (TPR * LC::XRS);
  const int64_t pr = (int64_t(blockIdx.x) * 256 + threadIdx.x) / TPR;
  int64_t total = npairs;
  if constexpr (LIST) {
    total = int64_t(*rv.nlive) * npairs;
    if (int64_t(blockIdx.x) * (256 / TPR) >= total) return;  // the block's slots are all empty
  }
  const bool live = pr < total;
  const int64_t row0 = live ? pr * 2 : 0;  // row pairs never straddle images: rows per image are even
  const float* a0 = A + row0 * W;
  if constexpr (LIST) {
    const int64_t slot = live ? pr / npairs : 0;
    a0 = run_state(rv, rv.list[slot], false) + (row0 - slot * npairs * 2) * W;
  }
  const float* a1 = a0 + W;
  float2 x[P];
#pragma unroll
  for (int m = 0; m < P; ++m) x[m] = make_float2(a0[TPR * m + t], a1[TPR * m + t]);
  fft_group<W, false>(x, t, lane, xch);
  wave_lds_fence();
#pragma unroll
  for (int e = 0; e < P; ++e) xch[freq_of<W>(t, e)] = x[e];
  wave_lds_fence();
  float2* s0 = S + row0 * int64_t(WH);
  float2* s1 = s0 + WH;
#pragma unroll
  for (int i = 0; i < (WH + TPR - 1) / TPR; ++i) {
    const int k = t + TPR * i;
    if (k < WH) {
      const float2 z = xch[k];
      const float2 zn = xch[k == 0 ? 0 : W - k];
      if (live) {
        s0[k] = make_float2(0.5f * (z.x + zn.x), 0.5f * (z.y - zn.y));
        s1[k] = make_float2(0.5f * (z.y + zn.y), 0.5f * (zn.x - z.x));
      }
    }
  }
}

// ---------------------------------------------------------------- lines along a leading axis
struct LineArgs {
  const float2* S;
  float2* U;
  const float2* Kh;
  int64_t k_bstride;
  int64_t nlines;      // images * A * WH
  int WH, A;
  int64_t a_stride, st, img_elems;
  int C, K;            // images per world (channels) and kernels (mode 0)
  int cin[LX_MAX_K];   // mode 0: kernel k reads channel cin[k]
  int mode;            // 0: conv, 1: forward in place, 2: inverse S -> U
  int conjk;           // mode 0: multiply by conj(K) (adjoint of the potential)
  // mode 0: Kh holds real quadrants (float [world][K][L/2 + 1][WH], 2-D): the spectrum of kernel
  // k at frequency f of a line is the real value at row min(f, L - f); k_elems: the elements (float2,
  // or float for quadrants) of one kernel's spectrum
  int kquad;
  int64_t k_elems;
  // mode 0, optional: the forward spectra of the input images, FA [img][half]
  float2* FA;
  // mode 0, optional: dspec[b][k] += X_k * conj(FAin[b][acc_cin[k]]) (kernel-spectrum gradient)
  const float2* FAin;
  float2* dspec;
  int C_fa;
  int acc_cin[LX_MAX_K];
  // batched run (LIST kernels): nlines is the lines of one slot's slot_imgs images; the spectra
  // of a slot's lines are those of world list[slot]
  const int32_t* list;
  const int32_t* nlive;
  int xcd;  // blocks in XCD-contiguous order (xcd_block) or in launch order
};

LineArgs line_args() {
  LineArgs a;
  a.S = nullptr; a.U = nullptr; a.Kh = nullptr; a.k_bstride = 0; a.nlines = 0;
  a.WH = 0; a.A = 1; a.a_stride = 0; a.st = 0; a.img_elems = 0; a.C = 1; a.K = 0;
  for (int k = 0; k < LX_MAX_K; ++k) { a.cin[k] = -1; a.acc_cin[k] = 0; }
  a.mode = 0; a.conjk = 0; a.kquad = 0; a.k_elems = 0; a.FA = nullptr; a.FAin = nullptr; a.dspec = nullptr; a.C_fa = 1;
  a.list = nullptr; a.nlive = nullptr; a.xcd = 1;
  return a;
}

// One launch per mode (MODE 0: conv, 1: forward in place, 2: inverse), so that each keeps
// only its own registers.  Offsets inside an image are 32-bit; loads are issued in batches
// of 8 (sched_barrier) so that addresses and values are not all held at once.
// KQ (mode 0): the kernels' spectra are real quadrants (a separate instantiation, so that the
// complex layout's register allocation is its own)
template <int L, int MODE, bool LIST = false, int T = line_tpr(L, MODE), bool KQ = false>
__global__ __launch_bounds__(256) void lines_kernel(LineArgs a) {
  using LC = Lanes<L, T>;
  constexpr int TPR = LC::TPR, P = LC::P;
  __shared__ __align__(16) float2 xch_all[LC::XL ? LC::XCH : 2];
  const int lane = threadIdx.x & 63;
  const int t = threadIdx.x % TPR;
  float2* xch = xch_all + (LC::XL ? (threadIdx.x / TPR) * (TPR * LC::XRS) : 0);
  const int64_t blk = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : int64_t(blockIdx.x);
  const int64_t lid = (blk * 256 + threadIdx.x) / TPR;
  int64_t nlines = a.nlines;
  if constexpr (LIST) {
    nlines *= *a.nlive;
    if (blk * (256 / TPR) >= nlines) return;  // the block's slots are all empty
  }
  const bool live = lid < nlines;
  const int64_t per_img = int64_t(a.A) * a.WH;
  const int64_t img = live ? lid / per_img : 0;
  const int rem = live ? int(lid - img * per_img) : 0;
  const int aa = rem / a.WH;
  const int off = aa * int(a.a_stride) + (rem - aa * a.WH);  // line base inside its image
  const int64_t b = img / a.C;
  const int64_t bw = LIST && MODE == 0 ? int64_t(a.list[b]) : b;  // the world whose spectra apply
  const int c = int(img - b * a.C);
  const int st = int(a.st);
  const float2* src = a.S + img * a.img_elems + off;
  float2 x[P], y[P];
  if constexpr (MODE == 2) {  // inverse only; input in frequency order
#pragma unroll
    for (int e = 0; e < P; ++e) {
      y[e] = src[freq_of<L, TPR>(t, e) * st];
      if ((e & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    }
    fft_group<L, true, TPR>(y, t, lane, xch);
    float2* dst = a.U + img * a.img_elems + off;
    if (live) {
#pragma unroll
      for (int m = 0; m < P; ++m) dst[(TPR * m + t) * st] = y[m];
    }
    return;
  }
#pragma unroll
  for (int m = 0; m < P; ++m) {
    x[m] = src[(TPR * m + t) * st];
    if ((m & 7) == 7) __builtin_amdgcn_sched_barrier(0);
  }
  fft_group<L, false, TPR>(x, t, lane, xch);
  if constexpr (MODE == 1) {  // forward only, in place
    float2* dst = const_cast<float2*>(src);
    if (live) {
#pragma unroll
      for (int e = 0; e < P; ++e) dst[freq_of<L, TPR>(t, e) * st] = x[e];
    }
    return;
  }
  if (a.FA && live) {
    float2* fa = a.FA + img * a.img_elems + off;
#pragma unroll
    for (int e = 0; e < P; ++e) fa[freq_of<L, TPR>(t, e) * st] = x[e];
  }
  const float csign = a.conjk ? -1.f : 1.f;  // multiply by K or by conj(K)
  for (int k = 0; k < a.K; ++k) {
    if (a.cin[k] != c) continue;
    // an opaque copy of the lane index: the per-kernel addresses and twiddle loads are then
    // recomputed in each iteration instead of being hoisted out of the loop and kept live
    int t = threadIdx.x % TPR;
    __asm__ volatile("" : "+v"(t));
    if (a.dspec && live) {
      const float2* fa = a.FAin + (b * a.C_fa + a.acc_cin[k]) * a.img_elems + off;
      float2* ds = a.dspec + (b * a.K + k) * a.img_elems + off;
#pragma unroll
      for (int e = 0; e < P; ++e) {
        const int o = freq_of<L, TPR>(t, e) * st;
        ds[o] = cadd(ds[o], cmulc(x[e], fa[o]));
        if ((e & 7) == 7) __builtin_amdgcn_sched_barrier(0);
      }
    }
    const int64_t kb = bw * a.k_bstride + int64_t(k) * a.k_elems + off;
    if constexpr (KQ) {
      const float* kq = reinterpret_cast<const float*>(a.Kh) + kb;
#pragma unroll
      for (int e = 0; e < P; ++e) {
        const int f = freq_of<L, TPR>(t, e);
        y[e] = cmul(x[e], make_float2(kq[(f <= L / 2 ? f : L - f) * st], 0.f));
        if ((e & 7) == 7) __builtin_amdgcn_sched_barrier(0);
      }
    } else {
      const float2* kk = a.Kh + kb;
#pragma unroll
      for (int e = 0; e < P; ++e) {
        const float2 kv = kk[freq_of<L, TPR>(t, e) * st];
        y[e] = cmul(x[e], make_float2(kv.x, csign * kv.y));
        if ((e & 7) == 7) __builtin_amdgcn_sched_barrier(0);
      }
    }
    fft_group<L, true, TPR>(y, t, lane, xch);
    float2* dst = a.U + (b * a.K + k) * a.img_elems + off;
    if (live) {
#pragma unroll
      for (int m = 0; m < P; ++m) dst[(TPR * m + t) * st] = y[m];
    }
  }
}

// The convolution pass of one channel and one kernel on a latency-bound grid (a few worlds): the
// line's kernel spectrum is loaded before the forward transform, so that its latency overlaps the
// transform instead of following it.  The same arithmetic as lines_kernel's mode 0.
template <int L, int T, bool KQ = false>
__global__ __launch_bounds__(256) void lines_conv1_kernel(LineArgs a) {
  using LC = Lanes<L, T>;
  constexpr int TPR = LC::TPR, P = LC::P;
  __shared__ __align__(16) float2 xch_all[LC::XL ? LC::XCH : 2];
  const int lane = threadIdx.x & 63;
  const int t = threadIdx.x % TPR;
  float2* xch = xch_all + (LC::XL ? (threadIdx.x / TPR) * (TPR * LC::XRS) : 0);
  const int64_t blk = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : int64_t(blockIdx.x);
  const int64_t lid = (blk * 256 + threadIdx.x) / TPR;
  const bool live = lid < a.nlines;
  const int64_t per_img = int64_t(a.A) * a.WH;
  const int64_t img = live ? lid / per_img : 0;  // one channel: image = world
  const int rem = live ? int(lid - img * per_img) : 0;
  const int aa = rem / a.WH;
  const int off = aa * int(a.a_stride) + (rem - aa * a.WH);
  const int st = int(a.st);
  const float2* src = a.S + img * a.img_elems + off;
  float2 x[P], kv[P];
#pragma unroll
  for (int m = 0; m < P; ++m) x[m] = src[(TPR * m + t) * st];
  if constexpr (KQ) {
    const float* kq = reinterpret_cast<const float*>(a.Kh) + img * a.k_bstride + off;
#pragma unroll
    for (int e = 0; e < P; ++e) {
      const int f = freq_of<L, T>(t, e);
      kv[e] = make_float2(kq[(f <= L / 2 ? f : L - f) * st], 0.f);
    }
  } else {
    const float2* kk = a.Kh + img * a.k_bstride + off;
#pragma unroll
    for (int e = 0; e < P; ++e) kv[e] = kk[freq_of<L, T>(t, e) * st];
  }
  fft_group<L, false, T>(x, t, lane, xch);
  if (a.FA && live) {
    float2* fa = a.FA + img * a.img_elems + off;
#pragma unroll
    for (int e = 0; e < P; ++e) fa[freq_of<L, T>(t, e) * st] = x[e];
  }
  const float csign = a.conjk ? -1.f : 1.f;
#pragma unroll
  for (int e = 0; e < P; ++e) x[e] = cmul(x[e], make_float2(kv[e].x, csign * kv[e].y));
  fft_group<L, true, T>(x, t, lane, xch);
  float2* dst = a.U + img * a.img_elems + off;
  if (live) {
#pragma unroll
    for (int m = 0; m < P; ++m) dst[(TPR * m + t) * st] = x[m];
  }
}

// The convolution pass of the adjoint potential with the kernel-spectrum gradient, one channel and
// one kernel (the reverse steps of lx_adjoint_step and of a run): the line's kernel spectrum, state
// spectrum and gradient accumulator are loaded before the forward transform, so that their latency
// overlaps it instead of following it (the single-world gradient is bound by these chains).  The
// same arithmetic as lines_kernel's mode 0 with dspec.
template <int L, int T>
__global__ __launch_bounds__(256) void lines_gradk_kernel(LineArgs a) {
  using LC = Lanes<L, T>;
  constexpr int TPR = LC::TPR, P = LC::P;
  __shared__ __align__(16) float2 xch_all[LC::XL ? LC::XCH : 2];
  const int lane = threadIdx.x & 63;
  const int t = threadIdx.x % TPR;
  float2* xch = xch_all + (LC::XL ? (threadIdx.x / TPR) * (TPR * LC::XRS) : 0);
  const int64_t blk = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : int64_t(blockIdx.x);
  const int64_t lid = (blk * 256 + threadIdx.x) / TPR;
  const bool live = lid < a.nlines;
  const int64_t per_img = int64_t(a.A) * a.WH;
  const int64_t img = live ? lid / per_img : 0;  // one channel: image = world
  const int rem = live ? int(lid - img * per_img) : 0;
  const int aa = rem / a.WH;
  const int off = aa * int(a.a_stride) + (rem - aa * a.WH);
  const int st = int(a.st);
  const float2* src = a.S + img * a.img_elems + off;
  const float2* kk = a.Kh + img * a.k_bstride + off;
  const float2* fa = a.FAin + img * a.img_elems + off;
  float2* ds = a.dspec + img * a.img_elems + off;
  float2 x[P], kv[P], fv[P], dv[P];
#pragma unroll
  for (int m = 0; m < P; ++m) x[m] = src[(TPR * m + t) * st];
#pragma unroll
  for (int e = 0; e < P; ++e) {
    const int o = freq_of<L, T>(t, e) * st;
    kv[e] = kk[o];
    fv[e] = fa[o];
    dv[e] = ds[o];
  }
  fft_group<L, false, T>(x, t, lane, xch);
  const float csign = a.conjk ? -1.f : 1.f;
#pragma unroll
  for (int e = 0; e < P; ++e) {
    if (live) ds[freq_of<L, T>(t, e) * st] = cadd(dv[e], cmulc(x[e], fv[e]));
    x[e] = cmul(x[e], make_float2(kv[e].x, csign * kv[e].y));
  }
  fft_group<L, true, T>(x, t, lane, xch);
  float2* dst = a.U + img * a.img_elems + off;
  if (live) {
#pragma unroll
    for (int m = 0; m < P; ++m) dst[(TPR * m + t) * st] = x[m];
  }
}

// The convolution pass (mode 0 without FA or dspec) with its global reads and writes staged
// through LDS.  A block holds LB = 256 / TPR lines; neighbouring lines are neighbouring half-
// spectrum columns, so a load or store instruction of the staging loop covers LB columns of
// 256 / LB points (16 x 8 B = 128 B per point at 16 lanes per line) where the group layout covers
// only 64 / TPR columns, which left the pass bound by its address unit.  Thread i moves line
// i % LB, points i / LB + TPR q; group g reads line g from the stage in its transform layout.
// Stage row g lies inside group g's exchange scratch (offset 2g points, which rotates the rows'
// banks), so a group's transforms touch only its own row: block barriers separate the staging
// steps (rows written or read by other groups' threads), wave-scope fences the rest.
// Used at 256 points (16 lanes per line: 3C6K step of 256 worlds of 256^2 0.745 -> 0.695 ms);
// at 96 points (8 lanes per line, 8 columns per instruction already) the stage's registers
// cost a wave per SIMD and the 96^3 adjoint ran 1.5% slower, so other lengths keep lines_kernel.
constexpr bool conv_staged(int L) { return L == 256; }
// ... and only for grids of at least 4 blocks per CU: a small grid is latency-bound, and the
// stage's block barriers serialise each block's loads (one 256^2 world, 9 blocks: 64 forward
// steps 1.63 -> 2.05 ms staged)
constexpr unsigned kStagedMinBlocks = 1024;

template <int L, bool LIST = false, bool KQ = false>
__global__ __launch_bounds__(256) void conv_staged_kernel(LineArgs a) {
  using LC = Lanes<L, line_tpr(L, 0)>;
  constexpr int TPR = LC::TPR, P = LC::P, LB = 256 / TPR, GRS = TPR * LC::XRS;
  static_assert(LC::XL && L + 2 * (LB - 1) <= GRS, "stage rows fit their groups' exchange scratch");
  __shared__ __align__(16) float2 xch_all[LC::XCH];
  __shared__ unsigned chan_mask;
  const int tid = threadIdx.x, lane = tid & 63, t = tid % TPR, g = tid / TPR;
  float2* xch = xch_all + g * GRS;
  auto srow = [](int r) { return r * (GRS + 2); };  // stage row r: r GRS + 2 r
  const int64_t blk = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : int64_t(blockIdx.x);
  int64_t nlines = a.nlines;
  if constexpr (LIST) {
    nlines *= *a.nlive;
    if (blk * LB >= nlines) return;  // the block's slots are all empty (uniform: before any barrier)
  }
  const int64_t per_img = int64_t(a.A) * a.WH;
  const int st = int(a.st);
  // the line this thread stages (ls) and its first point (pl)
  const int ls = tid % LB, pl = tid / LB;
  const int64_t lid_s = blk * LB + ls;
  const bool live_s = lid_s < nlines;
  const int64_t img_s = live_s ? lid_s / per_img : 0;
  const int rem_s = live_s ? int(lid_s - img_s * per_img) : 0;
  const int aa_s = rem_s / a.WH;
  const int off_s = aa_s * int(a.a_stride) + (rem_s - aa_s * a.WH);
  const int64_t b_s = img_s / a.C;
  const int64_t bw_s = LIST ? int64_t(a.list[b_s]) : b_s;  // the world whose spectra apply
  const int c_s = int(img_s - b_s * a.C);
  // the line this thread's group transforms
  const int64_t lid = blk * LB + g;
  const bool live = lid < nlines;
  const int64_t img = live ? lid / per_img : 0;
  const int c = int(img - (img / a.C) * a.C);
  float2* stage = xch_all;

  if (tid == 0) chan_mask = 0u;
  {
    const float2* src = a.S + img_s * a.img_elems + off_s;
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int pt = pl + TPR * q;
      stage[srow(ls) + pt] = live_s ? src[pt * st] : make_float2(0.f, 0.f);
      if ((q & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    }
  }
  __syncthreads();
  if (live_s) atomicOr(&chan_mask, 1u << c_s);
  float2 x[P], y[P];
#pragma unroll
  for (int m = 0; m < P; ++m) x[m] = stage[srow(g) + TPR * m + t];
  __syncthreads();
  const unsigned mask = chan_mask;  // channels of the block's lines (every OR precedes the barrier)
  fft_group<L, false, TPR>(x, t, lane, xch);
  const float csign = a.conjk ? -1.f : 1.f;  // multiply by K or by conj(K)
  for (int k = 0; k < a.K; ++k) {
    const int ck = a.cin[k];
    if (!((mask >> ck) & 1u)) continue;  // no line of the block reads kernel k (block-uniform)
    const bool use_s = live_s && c_s == ck;
    __syncthreads();  // the previous transform or store step is done with the scratch
    const int64_t kb = bw_s * a.k_bstride + int64_t(k) * a.k_elems + off_s;
    if constexpr (KQ) {
      // points 0..L/2 of the quadrant's column, each staged at f and L - f: points pl + TPR q
      // (q < P/2) are below L/2 for every thread, so their loads issue together; L/2 is thread
      // pl = 0's
      static_assert(P % 2 == 0 && TPR * (P / 2) == L / 2, "half the points below L/2");
      const float* kq = reinterpret_cast<const float*>(a.Kh) + kb;
      float v[P / 2 + 1];
#pragma unroll
      for (int q = 0; q < P / 2; ++q) v[q] = use_s ? kq[(pl + TPR * q) * st] : 0.f;
      v[P / 2] = (use_s && pl == 0) ? kq[(L / 2) * st] : 0.f;
      if (use_s) {
#pragma unroll
        for (int q = 0; q < P / 2; ++q) {
          const int pt = pl + TPR * q;
          stage[srow(ls) + pt] = make_float2(v[q], 0.f);
          if (pt > 0) stage[srow(ls) + L - pt] = make_float2(v[q], 0.f);
        }
        if (pl == 0) stage[srow(ls) + L / 2] = make_float2(v[P / 2], 0.f);
      }
    } else {
      const float2* kk = a.Kh + kb;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int pt = pl + TPR * q;
        if (use_s) stage[srow(ls) + pt] = kk[pt * st];
        if ((q & 7) == 7) __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();
    const bool use = live && c == ck;
#pragma unroll
    for (int e = 0; e < P; ++e) {
      const float2 kv = use ? stage[srow(g) + freq_of<L, TPR>(t, e)] : make_float2(0.f, 0.f);
      y[e] = cmul(x[e], make_float2(kv.x, csign * kv.y));
    }
    wave_lds_fence();  // the group's own row: read before its transform overwrites it
    fft_group<L, true, TPR>(y, t, lane, xch);
    wave_lds_fence();
#pragma unroll
    for (int m = 0; m < P; ++m) stage[srow(g) + TPR * m + t] = y[m];
    __syncthreads();
    if (use_s) {
      float2* dst = a.U + (b_s * a.K + k) * a.img_elems + off_s;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int pt = pl + TPR * q;
        dst[pt * st] = stage[srow(ls) + pt];
      }
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
  int64_t npairs;  // images * rows / 2
  int64_t rows;    // rows per image
  int C, K;
  int gfk[LX_MAX_K];
  int state_fn, average;
  int update;      // 0: potential only
  // selection mode: weight of kernel k in channel c is (cin[k] == c), growth is the identity and the
  // update is A_out = A + G (no gf / w / dt reads).  Used for the potential alone, the adjoint of the
  // potential (lambda += correlations) and the inverse real FFT.
  int sel;
  int cin[LX_MAX_K];
  // statistics mode: the step's world sums of leniax/statistics.py:35-126 over each row pair's
  // input state, field and potential, LX_PAIR_REC floats per pair at part[pair index] (sA, sG,
  // nA, nG, nU, sAy[3], sGy[3], sAyy[3]), in the re-centred coordinates of the carried shift
  float* part;
  const lx_stat_carry* carry;
  int ndim, d0, d1;  // rows = d0 d1 (d0 = 1 for 2-D worlds)
  // batched run (LIST kernels): npairs is the pairs of one slot; state from / to rv's buffers, the
  // carry is rv.carry; with hist_B > 0, field and pot (and cells_hist, when set) are the run's
  // histories [max_steps][hist_B][C or K][dims], written at the world's step
  RunView rv;
  float* cells_hist;
  int64_t hist_B;
  // the reverse step's gradient scalars, summed by fin.blocks workgroups after the row pairs' (not in
  // the statistics or LIST kernels)
  LxAdjFinish fin;
};

// UNI: every wave's row pairs lie in one image (the host checks that the pairs per image are a
// multiple of the pairs per wave), so the world's weights and growth parameters are the same for
// the whole wave: the wave copies them into LDS once instead of loading them from global memory
// in each kernel iteration, and its branches on them are uniform.
// PF (latency-bound grids, the last worlds of a batched run): the rows of the next kernel the loop
// transforms are loaded into registers while it transforms the current one (36 more VGPRs at 256
// points); the same arithmetic, in the same order.
template <int W, bool STATS, bool UNI, bool LIST = false, bool PF = false>
__global__ __launch_bounds__(256) void rows_inv_kernel(RowArgs a) {
  using LC = Lanes<W>;
  constexpr int TPR = LC::TPR, P = LC::P, WH = W / 2 + 1;
  __shared__ __align__(16) float2 xch_all[LC::XCH];
  __shared__ float wpar[UNI ? 4 : 1][UNI ? LX_MAX_C * LX_MAX_K + 2 * LX_MAX_K : 1];
  const int lane = threadIdx.x & 63;
  const int t = threadIdx.x % TPR;
  float2* xch = xch_all + (threadIdx.x / TPR) * (TPR * LC::XRS);
  if constexpr (!STATS && !LIST) {
    if (a.fin.blocks > 0 && int(blockIdx.x) >= int(gridDim.x) - a.fin.blocks) {
      const int64_t i = int64_t(int(blockIdx.x) - (int(gridDim.x) - a.fin.blocks)) * 4 + (threadIdx.x >> 6);
      if (i < lx_adj_finish_outputs(a.fin)) lx_adj_finish_wave(a.fin, i);
      return;
    }
  }
  const int64_t pr = (int64_t(blockIdx.x) * 256 + threadIdx.x) / TPR;
  int64_t npairs = a.npairs;
  if constexpr (LIST) {
    npairs *= *a.rv.nlive;
    if (int64_t(blockIdx.x) * (256 / TPR) >= npairs) return;  // the block's slots are all empty
  }
  const bool live = pr < npairs;
  const int64_t half_rows = a.rows / 2;
  const int64_t img = live ? pr / half_rows : 0;  // image of the pair (LIST: of its slot)
  const int64_t r2 = live ? (pr - img * half_rows) * 2 : 0;
  int64_t b = img / a.C;  // world (LIST: slot)
  int c = int(img - b * a.C);
  int64_t bw = b;         // world
  if constexpr (LIST) bw = a.rv.list[b];
  const float inv_n = 1.f / float(a.rows * W);
  const float* pw = a.sel ? nullptr : a.w + bw * a.w_bs;
  const float* pg = a.sel ? nullptr : a.gf + bw * a.gf_bs;
  if constexpr (UNI) {
    b = __builtin_amdgcn_readfirstlane(int(b));
    bw = LIST ? int64_t(__builtin_amdgcn_readfirstlane(int(bw))) : b;
    c = __builtin_amdgcn_readfirstlane(c);
    if (!a.sel) {
      float* ws = wpar[threadIdx.x >> 6];
      const int nw = a.C * a.K;
      for (int i = lane; i < nw + 2 * a.K; i += 64)
        ws[i] = i < nw ? a.w[bw * a.w_bs + i] : a.gf[bw * a.gf_bs + (i - nw)];
      wave_lds_fence();
      pw = ws;
      pg = ws + nw;
    }
  }
  // the world's step (LIST: history rows and the state buffers follow it)
  const int64_t tw = LIST ? int64_t(a.rv.carry[bw].t) : 0;
  float G0[P], G1[P];
#pragma unroll
  for (int m = 0; m < P; ++m) G0[m] = G1[m] = 0.f;
  float den = 0.f;
  float nU = 0.f;  // statistics: potential cells above eps, over the kernels this channel owns
  float2 x[P];
  constexpr int NI = (WH + TPR - 1) / TPR;  // staged points of a row per lane
  float2 pf0[PF ? NI : 1], pf1[PF ? NI : 1];
  uint32_t act = 0;  // PF: the kernels the loop transforms (it skips the others)
  auto load_rows = [&](int k) {
    const float2* u0 = a.U + ((b * a.K + k) * a.rows + r2) * WH;
    const float2* u1 = u0 + WH;
    const int tl = threadIdx.x % TPR;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int col = tl + TPR * i;
      if (col < WH) {
        pf0[i] = u0[col];
        pf1[i] = u1[col];
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // the loads are issued here, before the transform that follows
  };
  // PF: the pair's input state, loaded before the transforms as well
  float av0[PF ? P : 1], av1[PF ? P : 1];
  if constexpr (PF) {
    if (a.update && live) {
      const float* Ap = LIST ? run_state(a.rv, bw, false) + (int64_t(c) * a.rows + r2) * W
                             : a.A + (img * a.rows + r2) * int64_t(W);
      const int tl = threadIdx.x % TPR;
#pragma unroll
      for (int m = 0; m < P; ++m) {
        av0[m] = Ap[TPR * m + tl];
        av1[m] = Ap[W + TPR * m + tl];
      }
    }
    for (int k = 0; k < a.K; ++k) {
      const float wk = a.sel ? (a.cin[k] == c ? 1.f : 0.f) : pw[c * a.K + k];
      int owner = 0;
      if (a.sel) {
        owner = a.cin[k];
      } else {
        for (int cc = a.C - 1; cc >= 0; --cc)
          if (pw[cc * a.K + k] != 0.f) owner = cc;
      }
      const bool own = owner == c;
      if (!(wk == 0.f && !(a.pot != nullptr && own) && !(STATS && own))) act |= 1u << k;
    }
    if (act) load_rows(__builtin_ctz(act));
  }
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
    // every kernel has one owner channel: it writes the kernel's potential and counts its cells
    const bool own = owner == c;
    const bool want_pot = a.pot != nullptr && own;
    if (wk == 0.f && !want_pot && !(STATS && own)) continue;
    int t = threadIdx.x % TPR;  // opaque per iteration: addresses and twiddles are not hoisted
    __asm__ volatile("" : "+v"(t));
    // stage the two rows' half spectra (coalesced), then repack them as one complex row:
    // Z[k] = A[k] + i B[k], A[k] = conj(A[W-k]) above W/2; DC and Nyquist are real
    wave_lds_fence();
    if constexpr (PF) {
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int col = t + TPR * i;
        if (col < WH) {
          xch[col] = pf0[i];
          xch[WH + col] = pf1[i];
        }
      }
      const uint32_t rest = act & ~((2u << k) - 1u);
      if (rest) load_rows(__builtin_ctz(rest));
    } else {
      const float2* u0 = a.U + ((b * a.K + k) * a.rows + r2) * WH;
      const float2* u1 = u0 + WH;
#pragma unroll
      for (int i = 0; i < (WH + TPR - 1) / TPR; ++i) {
        const int col = t + TPR * i;
        if (col < WH) {
          xch[col] = u0[col];
          xch[WH + col] = u1[col];
        }
      }
    }
    wave_lds_fence();
#pragma unroll
    for (int e = 0; e < P; ++e) {
      const int kk = freq_of<W>(t, e);
      const bool mirror = kk >= W / 2;
      const float sg = mirror ? -1.f : 1.f;
      const int col = mirror ? W - kk : kk;
      const float2 ya = xch[col], yb = xch[WH + col];
      float2 z = make_float2(fmaf(-sg, yb.y, ya.x), fmaf(sg, ya.y, yb.x));
      if (kk == 0 || kk == W / 2) z = make_float2(ya.x, yb.x);
      x[e] = z;
    }
    wave_lds_fence();
    fft_group<W, true>(x, t, lane, xch);
    const float gm = a.sel ? 0.f : pg[2 * k], gs = a.sel ? 1.f : pg[2 * k + 1];
    const float inv_s = 1.f / gs, inv_9s2 = 1.f / (9.f * (gs * gs));
    float* po0 = !want_pot ? nullptr
                 : LIST ? a.pot + (((tw * a.hist_B + bw) * a.K + k) * a.rows + r2) * W
                        : a.pot + ((b * a.K + k) * a.rows + r2) * W;
    // the growth function is chosen once per kernel, outside the loop over the lane's points
    // (KIND < 0: selection mode, identity)
    auto accumulate = [&](auto kind_c) {
      constexpr int KIND = decltype(kind_c)::value;
#pragma unroll
      for (int m = 0; m < P; ++m) {
        const float U0 = x[m].x * inv_n, U1 = x[m].y * inv_n;
        if (want_pot && live) {
          po0[TPR * m + t] = U0;
          po0[W + TPR * m + t] = U1;
        }
        if (a.update) {
          float g0 = U0, g1 = U1;
          if constexpr (KIND >= 0) {
            g0 = growth_pre<KIND>(gm, gs, inv_s, inv_9s2, U0);
            g1 = growth_pre<KIND>(gm, gs, inv_s, inv_9s2, U1);
          }
          G0[m] = add_nc(G0[m], mul_nc(wk, g0));
          G1[m] = add_nc(G1[m], mul_nc(wk, g1));
        }
      }
    };
    // the count has a loop of its own: folded into the growth loop below, it keeps the kernel
    // at 156 VGPRs (3 waves per SIMD) instead of 128 for rows of 64 and 128 points
    if constexpr (STATS) {
      if (own) {
#pragma unroll
        for (int m = 0; m < P; ++m)
          nU += (x[m].x * inv_n > LX_EPS ? 1.f : 0.f) + (x[m].y * inv_n > LX_EPS ? 1.f : 0.f);
      }
    }
    if (a.sel) {
      accumulate(IntC<-1>{});
    } else {
      switch (a.gfk[k]) {
        case LX_GF_POLY_QUAD4: accumulate(IntC<LX_GF_POLY_QUAD4>{}); break;
        case LX_GF_GAUSSIAN: accumulate(IntC<LX_GF_GAUSSIAN>{}); break;
        case LX_GF_GAUSSIAN_TARGET: accumulate(IntC<LX_GF_GAUSSIAN_TARGET>{}); break;
        case LX_GF_STEP: accumulate(IntC<LX_GF_STEP>{}); break;
        case LX_GF_STAIRCASE: accumulate(IntC<LX_GF_STAIRCASE>{}); break;
        case LX_GF_TRIANGLE: accumulate(IntC<LX_GF_TRIANGLE>{}); break;
        default: accumulate(IntC<LX_GF_IDENTITY>{}); break;
      }
    }
  }
  if (!a.update || !live) return;
  const float dtw = a.sel ? 1.f : a.dt[bw * a.dt_bs];
  const int64_t off = (img * a.rows + r2) * int64_t(W);
  const float* Ain = a.A + off;
  float* Aout = a.A_out + off;
  float* fld = a.field ? a.field + off : nullptr;
  float* chist = nullptr;
  if constexpr (LIST) {
    const int64_t in_world = (int64_t(c) * a.rows + r2) * W;
    Ain = run_state(a.rv, bw, false) + in_world;
    Aout = run_state(a.rv, bw, true) + in_world;
    const int64_t hist_off = (tw * a.hist_B + bw) * a.rv.world_elems + in_world;
    fld = a.field ? a.field + hist_off : nullptr;
    chist = a.cells_hist ? a.cells_hist + hist_off : nullptr;
  }
  // statistics: per row sums of A, A z, A z^2, G+, G+ z (z the re-centred column) and counts
  float sAr0 = 0.f, sAr1 = 0.f, sAz = 0.f, sAzz = 0.f, sGr0 = 0.f, sGr1 = 0.f, sGz = 0.f, nA = 0.f, nG = 0.f;
  int sh[3] = {0, 0, 0};  // carried shift reduced into [0, size) along (d0, d1, W)
  if (STATS) {
    auto modp = [](int v, int m) { v %= m; return v < 0 ? v + m : v; };
    const lx_carry& cc = LIST ? a.rv.carry[bw] : a.carry[b].c;
    if (a.ndim == 3) sh[0] = modp(cc.shift[0], a.d0);
    sh[1] = modp(cc.shift[a.ndim - 2], a.d1);
    sh[2] = modp(cc.shift[a.ndim - 1], W);
  }
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int col = TPR * m + t;
    float g0 = G0[m], g1 = G1[m];
    if (a.average) { g0 = (g0 / den); g1 = (g1 / den); }
    if (fld) { fld[col] = g0; fld[W + col] = g1; }
    float v0, v1;
    if constexpr (PF) {
      v0 = av0[m];
      v1 = av1[m];
    } else {
      v0 = Ain[col];
      v1 = Ain[W + col];
    }
    if (LIST && chist) { chist[col] = v0; chist[W + col] = v1; }
    if (STATS) {
      int zc = col - sh[2];
      zc += zc < 0 ? W : 0;
      const float cz = float(zc - W / 2);
      const float gp0 = g0 != g0 ? g0 : fmaxf(g0, 0.f), gp1 = g1 != g1 ? g1 : fmaxf(g1, 0.f);
      sAr0 += v0;
      sAr1 += v1;
      const float vs = v0 + v1, gs2 = gp0 + gp1;
      sAz = fmaf(vs, cz, sAz);
      sAzz = fmaf(vs * cz, cz, sAzz);
      sGr0 += gp0;
      sGr1 += gp1;
      sGz = fmaf(gs2, cz, sGz);
      nA += (v0 > LX_EPS ? 1.f : 0.f) + (v1 > LX_EPS ? 1.f : 0.f);
      nG += (gp0 > LX_EPS ? 1.f : 0.f) + (gp1 > LX_EPS ? 1.f : 0.f);
    }
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
    Aout[col] = n0;
    Aout[W + col] = n1;
  }
  if constexpr (STATS) {
    // re-centred row coordinates of the pair's two rows: (x, y) in 3-D, y in 2-D
    float cx[2], cy[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int row = int(r2) + i;
      int x = row / a.d1, y = row - (row / a.d1) * a.d1;
      x -= sh[0]; x += x < 0 ? a.d0 : 0;
      y -= sh[1]; y += y < 0 ? a.d1 : 0;
      cx[i] = float(x - a.d0 / 2);
      cy[i] = float(y - a.d1 / 2);
    }
    const bool d3 = a.ndim == 3;
    float rec[LX_PAIR_REC];
    rec[0] = sAr0 + sAr1;
    rec[1] = sGr0 + sGr1;
    rec[2] = nA;
    rec[3] = nG;
    rec[4] = nU;
    const float ay = fmaf(sAr0, cy[0], sAr1 * cy[1]), ayy = fmaf(sAr0 * cy[0], cy[0], sAr1 * cy[1] * cy[1]);
    const float gy = fmaf(sGr0, cy[0], sGr1 * cy[1]);
    const float ax = fmaf(sAr0, cx[0], sAr1 * cx[1]), axx = fmaf(sAr0 * cx[0], cx[0], sAr1 * cx[1] * cx[1]);
    const float gx = fmaf(sGr0, cx[0], sGr1 * cx[1]);
    rec[5] = d3 ? ax : ay;     rec[6] = d3 ? ay : sAz;   rec[7] = d3 ? sAz : 0.f;
    rec[8] = d3 ? gx : gy;     rec[9] = d3 ? gy : sGz;   rec[10] = d3 ? sGz : 0.f;
    rec[11] = d3 ? axx : ayy;  rec[12] = d3 ? ayy : sAzz; rec[13] = d3 ? sAzz : 0.f;
    for (int i = 14; i < LX_PAIR_REC; ++i) rec[i] = 0.f;
    // sum over the pair's TPR lanes (a butterfly: every lane ends with the same sums), then
    // lane t stores fields t, t + TPR, ...
#pragma unroll
    for (int off2 = TPR / 2; off2 >= 1; off2 >>= 1)
#pragma unroll
      for (int i = 0; i < 14; ++i) rec[i] += __shfl_xor(rec[i], off2, 64);
    // the pair's index rebuilt from (img, r2), which stay live for the update, rather than pr
    // kept live over the kernel loop (two registers more at the loop's peak)
    float* dst = a.part + (img * half_rows + (r2 >> 1)) * LX_PAIR_REC;
#pragma unroll
    for (int e = 0; e < LX_PAIR_REC; ++e)
      if (e % TPR == t) dst[e] = rec[e];
  }
}

// ---------------------------------------------------------------- launchers
int launch_rows_fwd(const Geom& g, const float* A, float2* S, int64_t nimg, hipStream_t st) {
  const int64_t npairs = nimg * (g.rows / 2);
  return dispatch_len(g.W, [&](auto w) {
    constexpr int W = decltype(w)::value;
    if (int rc = ensure_tw<W>(st)) return rc;
    hipLaunchKernelGGL((rows_fwd_kernel<W, false>), dim3(blocks_for(npairs * Lanes<W>::TPR)), dim3(256), 0, st, A,
                       S, npairs, RunView{nullptr, nullptr, nullptr, nullptr, nullptr, 0});
    return launched("rows_fwd");
  });
}

// ---------------------------------------------------------------- line pass variants and their choice
// A line pass has schedule variants that compute the same transforms: the convolution pass at 256
// points staged through LDS or in the lane groups' layout, blocks in XCD-contiguous or launch order,
// and at 96 points 4 or 8 lanes per line.  The fixed rule (tuning off) is the measured default:
// staged from 1024 blocks up, XCD order, 4 lanes for the forward / inverse passes at 96 points.
struct LineVariant {
  int staged, xcd, tpr;
};

// launches one variant of a line pass over `lines` lines (LIST: the slots' lines of a batched run)
template <bool LIST>
int launch_lines_v(LineArgs la, int L, int64_t lines, LineVariant v, hipStream_t st) {
  la.xcd = v.xcd;
  return dispatch_len(L, [&](auto l) {
    constexpr int LL = decltype(l)::value;
    auto go = [&](auto tc) {
      constexpr int T = decltype(tc)::value;
      const dim3 grid(blocks_for(lines * T));
      if (la.mode == 0) {
        auto conv = [&](auto kq) {
          constexpr bool KQ = decltype(kq)::value;
          if constexpr (conv_staged(LL) && T == line_tpr(LL, 0)) {
            if (v.staged) {
              hipLaunchKernelGGL((conv_staged_kernel<LL, LIST, KQ>), grid, dim3(256), 0, st, la);
              return launched("conv (staged)");
            }
          }
          // latency-bound grids (less than two blocks per CU) only: its registers cost occupancy
          if constexpr (!KQ) {
            if (!LIST && la.dspec && !la.FA && la.C == 1 && la.K == 1 && la.C_fa == 1 && la.acc_cin[0] == 0 &&
                la.cin[0] == 0 && grid.x < 512) {
              hipLaunchKernelGGL((lines_gradk_kernel<LL, T>), grid, dim3(256), 0, st, la);
              return launched("lines (kernel gradient)");
            }
          }
          if (!LIST && !la.dspec && la.C == 1 && la.K == 1 && la.cin[0] == 0 && grid.x < 512) {
            hipLaunchKernelGGL((lines_conv1_kernel<LL, T, KQ>), grid, dim3(256), 0, st, la);
            return launched("lines (convolution, one kernel)");
          }
          hipLaunchKernelGGL((lines_kernel<LL, 0, LIST, T, KQ>), grid, dim3(256), 0, st, la);
          return launched("lines");
        };
        return la.kquad ? conv(std::true_type{}) : conv(std::false_type{});
      } else if (la.mode == 1) {
        hipLaunchKernelGGL((lines_kernel<LL, 1, LIST, T>), grid, dim3(256), 0, st, la);
      } else {
        hipLaunchKernelGGL((lines_kernel<LL, 2, LIST, T>), grid, dim3(256), 0, st, la);
      }
      return launched("lines");
    };
    if constexpr (LL == 96) {
      if (v.tpr == 4) return go(IntC<4>{});
    }
    return go(IntC<row_tpr(LL)>{});
  });
}

bool staged_eligible(const LineArgs& la, int L) { return la.mode == 0 && L == 256 && !la.FA && !la.dspec; }

LineVariant default_variant(const LineArgs& la, int L, int64_t lines) {
  LineVariant v;
  v.tpr = L == 96 ? (la.mode == 0 ? 8 : 4) : row_tpr(L);
  v.staged = staged_eligible(la, L) && blocks_for(lines * v.tpr) >= kStagedMinBlocks;
  v.xcd = 1;
  return v;
}

void lines_key(char* key, size_t n, const LineArgs& la, int L, int64_t lines, bool list) {
  snprintf(key, n, "L=%d:mode=%d:lines=%lld:K=%d:%s%s", L, la.mode, (long long)lines, la.mode == 0 ? la.K : 0,
           list ? "list" : "plain", la.mode == 0 && la.kquad ? ":quadrant" : "");
}

// The variant of a line pass: forced knobs, then a recorded decision, then (tuning on) the fastest
// candidate timed by run(v) -- which launches v on buffers it may overwrite -- then the fixed rule.
template <class Run>
LineVariant choose_lines(const LineArgs& la, int L, int64_t lines, bool list, bool can_time, hipStream_t st,
                         Run&& run) {
  char key[160];
  lines_key(key, sizeof(key), la, L, lines, list);
  LineVariant d = default_variant(la, L, lines), v = d;
  int fs = -1, fx = -1, ft = -1, val = 0;
  const bool hs = lx_tune_lookup("conv_staged", key, &val) != 0;
  if (hs) fs = val;
  const bool hx = lx_tune_lookup("line_xcd", key, &val) != 0;
  if (hx) fx = val;
  const bool ht = lx_tune_lookup("lines96_tpr", key, &val) != 0;
  if (ht) ft = val;
  if (hs) v.staged = staged_eligible(la, L) && fs != 0;
  if (hx) v.xcd = fx != 0;
  if (ht && L == 96) v.tpr = ft == 4 ? 4 : 8;
  const bool all_known = (hs || !staged_eligible(la, L)) && hx && (ht || L != 96);
  if (all_known || !can_time || !lx_tune_should_time(st)) return v;
  // time every candidate the known values leave open
  LineVariant best = v;
  float best_ms = -1.f;
  for (int s = 0; s <= 1; ++s) {
    if (s && !staged_eligible(la, L)) continue;
    if (hs && s != v.staged) continue;
    for (int x = 0; x <= 1; ++x) {
      if (hx && x != v.xcd) continue;
      for (int t : {4, 8, 16}) {
        if (L == 96 ? (t != 4 && t != 8) : t != row_tpr(L)) continue;
        if (ht && L == 96 && t != v.tpr) continue;
        if (s && t != line_tpr(L, 0)) continue;
        LineVariant c{s, x, t};
        const float ms = lx_time_ms(st, 3, [&] { return run(c); });
        if (ms >= 0.f && (best_ms < 0.f || ms < best_ms)) {
          best_ms = ms;
          best = c;
        }
      }
    }
  }
  if (best_ms < 0.f) return d;
  if (staged_eligible(la, L)) lx_tune_record("conv_staged", key, best.staged);
  lx_tune_record("line_xcd", key, best.xcd);
  if (L == 96) lx_tune_record("lines96_tpr", key, best.tpr);
  return best;
}

// la: mode-specific fields filled by the caller; geometry of `axis` and the line count set here
int launch_lines(const Geom& g, int axis, LineArgs la, int64_t nimg, hipStream_t st) {
  la.WH = g.WH; la.A = g.A[axis]; la.a_stride = g.a_stride[axis]; la.st = g.st[axis]; la.img_elems = g.half;
  la.k_elems = la.kquad ? int64_t(g.L[0] / 2 + 1) * g.WH : g.half;
  la.nlines = nimg * g.A[axis] * int64_t(g.WH);
  const int L = g.L[axis];
  if (int rc = dispatch_len(L, [&](auto l) { return ensure_tw<decltype(l)::value>(st); })) return rc;
  // timing is safe where the pass can run again on its own inputs: a convolution into another
  // buffer with no accumulated kernel gradient, an out-of-place inverse; in-place passes are timed
  // on a copy of their images
  const bool in_place = la.mode == 1 || la.S == la.U;
  const bool idem = !in_place && !la.dspec;
  float2* tmp = nullptr;
  auto run = [&](LineVariant v) {
    LineArgs t = la;
    if (tmp) { t.S = tmp; t.U = tmp; }
    return launch_lines_v<false>(t, L, la.nlines, v, st);
  };
  const bool can_time = idem || (in_place && la.mode != 0);
  if (can_time && !idem && lx_tune_should_time(st)) {
    const size_t bytes = size_t(nimg) * size_t(g.half) * sizeof(float2);
    char key[160];
    lines_key(key, sizeof(key), la, L, la.nlines, false);
    int val = 0;
    const bool known = lx_tune_lookup("line_xcd", key, &val) && (L != 96 || lx_tune_lookup("lines96_tpr", key, &val));
    if (!known && hipMalloc(reinterpret_cast<void**>(&tmp), bytes) == hipSuccess) {
      if (hipMemcpyAsync(tmp, la.S, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) {
        hipFree(tmp);
        tmp = nullptr;
      }
    } else {
      tmp = nullptr;
    }
  }
  const LineVariant v = choose_lines(la, L, la.nlines, false, can_time && (idem || tmp), st, run);
  if (tmp) {
    hipStreamSynchronize(st);
    hipFree(tmp);
  }
  return launch_lines_v<false>(la, L, la.nlines, v, st);
}

int launch_rows_inv(const Geom& g, RowArgs ra, int64_t nimg, hipStream_t st) {
  ra.rows = g.rows;
  ra.npairs = nimg * (g.rows / 2);
  return dispatch_len(g.W, [&](auto w) {
    constexpr int W = decltype(w)::value;
    if (int rc = ensure_tw<W>(st)) return rc;
    const dim3 grid(blocks_for(ra.npairs * Lanes<W>::TPR) + unsigned(ra.fin.blocks));
    const bool uni = (ra.rows / 2) % (64 / Lanes<W>::TPR) == 0;  // waves never straddle images
    if (ra.part && ra.fin.blocks) return lx_fail(LX_EINVAL, "rows_inv: statistics and gradient sums together");
    if (!ra.part && grid.x < 512) {  // latency-bound: loads issued ahead of the transforms
      if (uni) hipLaunchKernelGGL((rows_inv_kernel<W, false, true, false, true>), grid, dim3(256), 0, st, ra);
      else hipLaunchKernelGGL((rows_inv_kernel<W, false, false, false, true>), grid, dim3(256), 0, st, ra);
      return launched("rows_inv");
    }
    if (ra.part) {
      if (uni) hipLaunchKernelGGL((rows_inv_kernel<W, true, true>), grid, dim3(256), 0, st, ra);
      else hipLaunchKernelGGL((rows_inv_kernel<W, true, false>), grid, dim3(256), 0, st, ra);
    } else {
      if (uni) hipLaunchKernelGGL((rows_inv_kernel<W, false, true>), grid, dim3(256), 0, st, ra);
      else hipLaunchKernelGGL((rows_inv_kernel<W, false, false>), grid, dim3(256), 0, st, ra);
    }
    return launched("rows_inv");
  });
}

RowArgs row_args(const lx_rule* rule) {
  RowArgs ra;
  ra.U = nullptr; ra.A = nullptr; ra.A_out = nullptr; ra.field = nullptr; ra.pot = nullptr;
  ra.gf = nullptr; ra.gf_bs = 0; ra.w = nullptr; ra.w_bs = 0; ra.dt = nullptr; ra.dt_bs = 0;
  ra.npairs = 0; ra.rows = 0; ra.C = rule->C; ra.K = rule->K;
  for (int k = 0; k < LX_MAX_K; ++k) {
    ra.gfk[k] = k < rule->K ? rule->gf[k] : 0;
    ra.cin[k] = k < rule->K ? rule->cin[k] : -1;
  }
  ra.state_fn = rule->state_fn; ra.average = rule->average; ra.update = 0; ra.sel = 0;
  ra.part = nullptr; ra.carry = nullptr;
  ra.rv = RunView{nullptr, nullptr, nullptr, nullptr, nullptr, 0};
  ra.cells_hist = nullptr; ra.hist_B = 0;
  ra.fin = LxAdjFinish{};
  ra.ndim = rule->ndim;
  ra.d0 = rule->ndim == 3 ? rule->dims[0] : 1;
  ra.d1 = rule->dims[rule->ndim - 2];
  return ra;
}

// forward transform of every image along the leading axes (after rows_fwd), in place
int fwd_leading(const Geom& g, float2* X, int64_t nimg, hipStream_t st) {
  for (int ax = g.naxes - 1; ax >= 0; --ax) {
    LineArgs la = line_args();
    la.S = X; la.mode = 1;
    if (int rc = launch_lines(g, ax, la, nimg, st)) return rc;
  }
  return LX_OK;
}

// One leniax step (or, with sel, the potential alone).  A: [B][C][dims]; work: S then U.
int plain_step(const lx_rule* rule, const float* A, int64_t B, const void* Kh, int64_t k_bs, const float* gf,
               int64_t gf_bs, const float* w, int64_t w_bs, const float* dt, int64_t dt_bs, float* A_out,
               float* field, float* pot, void* work, hipStream_t st, int sel, float2* FA = nullptr,
               float* part = nullptr, const lx_stat_carry* carry = nullptr) {
  Geom g;
  if (int rc = make_geom(rule->ndim, rule->dims, g)) return rc;
  const int C = rule->C, K = rule->K;
  float2* S = static_cast<float2*>(work);
  float2* U = S + B * C * g.half;
  int rc = launch_rows_fwd(g, A, S, B * C, st);
  if (!rc && g.naxes == 2) {  // forward along H
    LineArgs la = line_args();
    la.S = S; la.mode = 1;
    rc = launch_lines(g, 1, la, B * C, st);
  }
  if (!rc) {  // the convolution axis
    LineArgs la = line_args();
    la.S = S; la.U = U; la.Kh = static_cast<const float2*>(Kh); la.k_bstride = k_bs; la.C = C; la.K = K;
    la.kquad = rule->k_quadrant; la.FA = FA;
    for (int k = 0; k < LX_MAX_K; ++k) la.cin[k] = k < K ? rule->cin[k] : -1;
    rc = launch_lines(g, 0, la, B * C, st);
  }
  if (!rc && g.naxes == 2) {  // inverse along H, in place
    LineArgs la = line_args();
    la.S = U; la.U = U; la.mode = 2;
    rc = launch_lines(g, 1, la, B * K, st);
  }
  if (rc) return rc;
  RowArgs ra = row_args(rule);
  ra.U = U; ra.A = A; ra.A_out = A_out; ra.field = field; ra.pot = pot;
  ra.gf = gf; ra.gf_bs = gf_bs; ra.w = w; ra.w_bs = w_bs; ra.dt = dt; ra.dt_bs = dt_bs;
  ra.update = A_out != nullptr; ra.sel = sel;
  ra.part = part; ra.carry = carry;
  return launch_rows_inv(g, ra, B * C, st);
}

// full spectrum from the half spectrum of a real array (Hermitian extension), times scale
__global__ void expand_half_kernel(const float2* __restrict__ half, float2* __restrict__ full, int64_t n, int D,
                                   int H, int W, float scale) {
  const int WH = W / 2 + 1;
  const int64_t total = n * D * int64_t(H) * W;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int kw = int(i % W);
    const int kh = int((i / W) % H);
    const int kd = int((i / (int64_t(W) * H)) % D);
    const int64_t img = i / (int64_t(D) * H * W);
    const float2* hb = half + img * int64_t(D) * H * WH;
    const float2 v = kw <= W / 2 ? hb[(int64_t(kd) * H + kh) * WH + kw]
                                 : cconj(hb[(int64_t((D - kd) % D) * H + (H - kh) % H) * WH + (W - kw)]);
    full[i] = cscale(v, scale);
  }
}

int half_to_full(const Geom& g, int64_t n, const float2* half, float2* full, float scale, hipStream_t st) {
  const int D = g.ndim == 3 ? g.L[0] : 1;
  const int H = g.ndim == 3 ? g.L[1] : g.L[0];
  const int64_t total = n * D * int64_t(H) * g.W;
  const unsigned grid = unsigned(std::min<int64_t>((total + 255) / 256, 65536));
  hipLaunchKernelGGL(expand_half_kernel, dim3(grid), dim3(256), 0, st, half, full, n, D, H, g.W, scale);
  return launched("half to full");
}

// numpy.fft.irfftn (scaled by 1 / prod(dims)) of n half spectra; X untouched; work: n * half
int irfft(const Geom& g, int64_t n, const float2* X, float* x, float2* work, hipStream_t st) {
  for (int ax = 0; ax < g.naxes; ++ax) {
    LineArgs la = line_args();
    la.S = ax == 0 ? X : work; la.U = work; la.mode = 2;
    if (int rc = launch_lines(g, ax, la, n, st)) return rc;
  }
  lx_rule one{};
  one.ndim = g.ndim; one.C = 1; one.K = 1; one.cin[0] = 0;
  RowArgs ra = row_args(&one);
  ra.U = work; ra.pot = x; ra.sel = 1;
  return launch_rows_inv(g, ra, n, st);
}

}  // namespace

// ====================================================================== entry points
extern "C" int lx_workspace_size(const lx_rule* rule, int64_t B, size_t* bytes) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!bytes || B < 0) return lx_fail(LX_EINVAL, "bad arguments");
  int64_t rows = 1;
  for (int i = 0; i < rule->ndim - 1; ++i) rows *= rule->dims[i];
  const int64_t half = rows * (rule->dims[rule->ndim - 1] / 2 + 1);
  // spectra S [B][C] and U [B][K], then lx_step_stats' row-pair sums [B][C][rows / 2]
  *bytes = size_t(B) * size_t(rule->C + rule->K) * size_t(half) * sizeof(float2) +
           size_t(B) * size_t(rule->C) * size_t(rows / 2) * LX_PAIR_REC * sizeof(float);
  return LX_OK;
}

extern "C" int lx_update(const lx_rule* rule, const float* state_d, int64_t B, const void* K_half_d,
                         int64_t k_bstride, const float* gf_d, int64_t gf_bstride, const float* w_d, int64_t w_bstride,
                         const float* dt_d, int64_t dt_bstride, float* state_out_d, float* field_d,
                         float* potential_d, void* work_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  Geom g;
  if ((rc = make_geom(rule->ndim, rule->dims, g))) return rc;
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;  // nothing to read or write: pointers may be NULL
  if (!state_d || !K_half_d || !gf_d || !w_d || !dt_d || !state_out_d || !work_d)
    return lx_fail(LX_EINVAL, "NULL pointer");
  if (state_out_d == state_d) return lx_fail(LX_EINVAL, "lx_update: state_out must not alias state");
  return plain_step(rule, state_d, B, K_half_d, k_bstride, gf_d, gf_bstride, w_d, w_bstride, dt_d, dt_bstride,
                    state_out_d, field_d, potential_d, work_d, (hipStream_t)stream, 0);
}

extern "C" int lx_update_with_spectra(const lx_rule* rule, const float* state_d, int64_t B, const void* K_half_d,
                                      int64_t k_bstride, const float* gf_d, int64_t gf_bstride, const float* w_d,
                                      int64_t w_bstride, const float* dt_d, int64_t dt_bstride, float* state_out_d,
                                      float* field_d, float* potential_d, void* state_spec_d, void* work_d,
                                      void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  Geom g;
  if ((rc = make_geom(rule->ndim, rule->dims, g))) return rc;
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  if (!state_d || !K_half_d || !gf_d || !w_d || !dt_d || !state_out_d || !state_spec_d || !work_d)
    return lx_fail(LX_EINVAL, "NULL pointer");
  if (state_out_d == state_d) return lx_fail(LX_EINVAL, "lx_update_with_spectra: state_out must not alias state");
  return plain_step(rule, state_d, B, K_half_d, k_bstride, gf_d, gf_bstride, w_d, w_bstride, dt_d, dt_bstride,
                    state_out_d, field_d, potential_d, work_d, (hipStream_t)stream, 0,
                    static_cast<float2*>(state_spec_d));
}

extern "C" int lx_step_stats(const lx_rule* rule, const float* state_d, int64_t B, const void* K_half_d,
                             int64_t k_bstride, const float* gf_d, int64_t gf_bstride, const float* w_d,
                             int64_t w_bstride, const float* dt_d, int64_t dt_bstride, float* state_out_d,
                             float* field_d, float* potential_d, float R, float T, void* carry_d, float* stats_d,
                             void* work_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  Geom g;
  if ((rc = make_geom(rule->ndim, rule->dims, g))) return rc;
  if (B < 0 || !(R > 0.f) || !(T > 0.f)) return lx_fail(LX_EINVAL, "bad B / R / T");
  if (B == 0) return LX_OK;
  if (!state_d || !K_half_d || !gf_d || !w_d || !dt_d || !state_out_d || !carry_d || !stats_d || !work_d)
    return lx_fail(LX_EINVAL, "NULL pointer");
  if (state_out_d == state_d) return lx_fail(LX_EINVAL, "lx_step_stats: state_out must not alias state");
  const int64_t pairs = g.rows / 2;
  float* part = reinterpret_cast<float*>(static_cast<float2*>(work_d) + B * (rule->C + rule->K) * g.half);
  hipStream_t st = (hipStream_t)stream;
  if ((rc = plain_step(rule, state_d, B, K_half_d, k_bstride, gf_d, gf_bstride, w_d, w_bstride, dt_d, dt_bstride,
                       state_out_d, field_d, potential_d, work_d, st, 0, nullptr, part,
                       static_cast<const lx_stat_carry*>(carry_d))))
    return rc;
  return lx_pair_stats_launch(rule, part, B, pairs, R, T, carry_d, stats_d, st);
}

extern "C" int lx_potential(const lx_rule* rule, const float* state_d, int64_t B, const void* K_half_d,
                            int64_t k_bstride, float* potential_d, void* work_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  Geom g;
  if ((rc = make_geom(rule->ndim, rule->dims, g))) return rc;
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  if (!state_d || !K_half_d || !potential_d || !work_d) return lx_fail(LX_EINVAL, "NULL pointer");
  // selection mode: kernel k's potential is written by the row pass of its input channel; no update
  return plain_step(rule, state_d, B, K_half_d, k_bstride, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, nullptr,
                    potential_d, work_d, (hipStream_t)stream, 1);
}

extern "C" int lx_rfft2(int32_t ndim, const int32_t* dims, int64_t n, const float* x_d, void* X_d, void* stream) {
  if (!dims || n < 0) return lx_fail(LX_EINVAL, "bad arguments");
  Geom g;
  if (int rc = make_geom(ndim, dims, g)) return rc;
  if (n == 0) return LX_OK;
  if (!x_d || !X_d) return lx_fail(LX_EINVAL, "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  float2* X = static_cast<float2*>(X_d);
  if (int rc = launch_rows_fwd(g, x_d, X, n, st)) return rc;
  return fwd_leading(g, X, n, st);
}

extern "C" int lx_irfft2(int32_t ndim, const int32_t* dims, int64_t n, const void* X_d, float* x_d, void* work_d,
                         void* stream) {
  if (!dims || n < 0) return lx_fail(LX_EINVAL, "bad arguments");
  Geom g;
  if (int rc = make_geom(ndim, dims, g)) return rc;
  if (n == 0) return LX_OK;
  if (!x_d || !X_d || !work_d) return lx_fail(LX_EINVAL, "NULL pointer");
  return irfft(g, n, static_cast<const float2*>(X_d), x_d, static_cast<float2*>(work_d), (hipStream_t)stream);
}

extern "C" int lx_half_to_full(int32_t ndim, const int32_t* dims, int64_t n, const void* half_d, void* full_d,
                               float scale, void* stream) {
  if (!dims || n < 0) return lx_fail(LX_EINVAL, "bad arguments");
  Geom g;
  if (int rc = make_geom(ndim, dims, g)) return rc;
  if (n == 0) return LX_OK;
  if (!half_d || !full_d) return lx_fail(LX_EINVAL, "NULL pointer");
  return half_to_full(g, n, static_cast<const float2*>(half_d), static_cast<float2*>(full_d), scale,
                      (hipStream_t)stream);
}

extern "C" int lx_kernel_spectrum(int32_t ndim, const int32_t* dims, int64_t n, const float* k_d, void* K_full_d,
                                  void* work_d, void* stream) {
  if (!dims || n < 0) return lx_fail(LX_EINVAL, "bad arguments");
  Geom g;
  if (int rc = make_geom(ndim, dims, g)) return rc;
  if (n == 0) return LX_OK;
  if (!k_d || !K_full_d || !work_d) return lx_fail(LX_EINVAL, "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  float2* work = static_cast<float2*>(work_d);
  if (int rc = launch_rows_fwd(g, k_d, work, n, st)) return rc;
  if (int rc = fwd_leading(g, work, n, st)) return rc;
  return half_to_full(g, n, work, static_cast<float2*>(K_full_d), 1.f, st);
}

// lambda[b][cin k] += correlation of dU_k with kernel k; dspec[b][k] += fft(dU_k) conj(FA[b][cin k])
extern "C" int lx_adjoint_potential(const lx_rule* rule, int64_t B, const void* K_half_d, int64_t k_bstride,
                                    const float* dpot_d, const void* state_spec_d, void* dspec_d, float* lambda_d,
                                    void* work_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc || (rc = lx_no_quadrant(rule, "lx_adjoint_potential"))) return rc;
  if ((state_spec_d == nullptr) != (dspec_d == nullptr))
    return lx_fail(LX_EINVAL, "state_spec and dspec are given together or not at all");
  Geom g;
  if ((rc = make_geom(rule->ndim, rule->dims, g))) return rc;
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  if (!K_half_d || !dpot_d || !lambda_d || !work_d) return lx_fail(LX_EINVAL, "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  const int K = rule->K;
  float2* X = static_cast<float2*>(work_d);
  // the K gradient images of a world are transformed as K "channels", kernel k on image k
  if ((rc = launch_rows_fwd(g, dpot_d, X, B * K, st))) return rc;
  if (g.naxes == 2) {
    LineArgs la = line_args();
    la.S = X; la.mode = 1;
    if ((rc = launch_lines(g, 1, la, B * K, st))) return rc;
  }
  {
    LineArgs la = line_args();
    la.S = X; la.U = X;  // each line is read whole before it is written back
    la.Kh = static_cast<const float2*>(K_half_d); la.k_bstride = k_bstride; la.C = K; la.K = K; la.conjk = 1;
    for (int k = 0; k < LX_MAX_K; ++k) {
      la.cin[k] = k < K ? k : -1;
      la.acc_cin[k] = k < K ? rule->cin[k] : 0;
    }
    la.FAin = static_cast<const float2*>(state_spec_d);
    la.dspec = static_cast<float2*>(dspec_d);
    la.C_fa = rule->C;
    if ((rc = launch_lines(g, 0, la, B * K, st))) return rc;
  }
  if (g.naxes == 2) {
    LineArgs la = line_args();
    la.S = X; la.U = X; la.mode = 2;
    if ((rc = launch_lines(g, 1, la, B * K, st))) return rc;
  }
  RowArgs ra = row_args(rule);
  ra.U = X; ra.A = lambda_d; ra.A_out = lambda_d;  // lambda_c += sum over kernels fed by c (in place, per element)
  ra.update = 1; ra.sel = 1; ra.average = 0; ra.state_fn = LX_STATE_SIMPLE;
  return launch_rows_inv(g, ra, B * rule->C, st);
}

// ====================================================================== lx_run, multi-pass path
// leniax/runner.py:167-215 (run_scan_mem_optimized) and :119-164 (run_scan) for every rule the
// step passes take, with the per-step statistics, the stop heuristics evaluated on the device every
// step (leniax/statistics.py:134-205) and the retirement of worlds at uneven times.
//
// A world that stops frees its slot; a device-side control kernel compacts the live list after each
// step and refills the free slots from a queue of worlds not yet started, so that at most `slots`
// worlds are in flight (their S, U and record scratch) while all B are resident.  The step's
// kernels run over the live list (the LIST variants above); blocks of empty slots return at once.
// G steps are captured as one HIP graph and replayed; the host reads the live count after each
// replay (one replay of lag) and, once the queue is drained and the count falls, replays a graph
// captured for fewer slots (a power-of-two fraction of `slots`), whose grids are that much smaller.
namespace {

// steps per captured graph (even: step i of a graph reads list i & 1).  The fixed rule: 8 steps when
// a step moves more than 2^22 cells (its launches are a small part of it), 32 below (small worlds
// or batches, where the graph's launches are what it saves); the knob "graph_steps" overrides it.
int steps_per_graph(int64_t cells_in_flight) {
  int v = 0;
  if (lx_tune_lookup("graph_steps", "run", &v) && v > 0) return v + (v & 1);
  return cells_in_flight > (int64_t(1) << 22) ? 8 : 32;
}

struct RunCtl {
  int32_t* list0;
  int32_t* list1;
  int32_t* nlive;  // [2]: live slots of list0, list1
  int32_t* head;   // next world of the queue
  int32_t* done;   // [slots]: the slot's world stopped this step
  lx_carry* carry; // [B]
};

struct PassesLayout {
  size_t X1, carry, work, list0, list1, done, ctl, total;
  size_t S, U, part;  // inside work
};

int64_t cells_per_world(const lx_rule* r) {
  int64_t n = 1;
  for (int i = 0; i < r->ndim; ++i) n *= r->dims[i];
  return n;
}

PassesLayout passes_layout(const lx_rule* r, const Geom& g, int64_t B, int64_t slots) {
  auto up = [](size_t v) { return (v + 255) / 256 * 256; };
  PassesLayout l;
  size_t o = 0;
  l.X1 = o;    o = up(o + size_t(B) * r->C * cells_per_world(r) * sizeof(float));
  l.carry = o; o = up(o + size_t(B) * sizeof(lx_carry));
  l.work = o;
  l.S = o;     o = up(o + size_t(slots) * r->C * g.half * sizeof(float2));
  l.U = o;     o = up(o + size_t(slots) * r->K * g.half * sizeof(float2));
  l.part = o;  o = up(o + size_t(slots) * r->C * (g.rows / 2) * LX_PAIR_REC * sizeof(float));
  l.list0 = o; o = up(o + size_t(slots) * 4);
  l.list1 = o; o = up(o + size_t(slots) * 4);
  l.done = o;  o = up(o + size_t(slots) * 4);
  l.ctl = o;   o = up(o + 64);
  l.total = o;
  return l;
}

// every world's carry cleared; the first slots take worlds 0, 1, ... and the queue starts after them
__global__ void run_start_kernel(RunCtl c, int64_t B, int slots) {
  const int64_t n = B > slots ? B : slots;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    if (i < B) {
      lx_carry z;
      for (int d = 0; d < 3; ++d) { z.shift[d] = 0; z.centroid[d] = 0.f; }
      z.angle = 0.f;
      for (int ch = 0; ch < LX_MAX_C; ++ch) z.init_cm[ch] = 0.f;
      z.prev_mass = z.prev_sign = 0.f;
      z.mono = z.vol = z.n = z.t = z.pad = 0;
      z.cont = 1;
      c.carry[i] = z;
    }
    if (i < slots) c.list0[i] = int32_t(i < B ? i : 0);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int32_t first = int32_t(B < slots ? B : slots);
    c.nlive[0] = first;
    c.nlive[1] = 0;
    *c.head = first;
  }
}

struct RunStatArgs {
  RunCtl c;
  int par;             // the list this step reads
  const float* part;   // [slots][C][pairs] row-pair records
  int64_t pairs;
  int C, ndim;
  int32_t dims[3];
  float R2, R, dt;
  int stop_policy, min_steps, max_steps;
  float cells_total;
  float* stats;        // [B][max_steps][11 + C] or NULL
  int32_t* n_out;      // [B] or NULL
  int32_t* steps_out;  // [B] or NULL
};

// One block per slot: the step's statistics of the slot's world (leniax/statistics.py:64-126), its
// carry, the heuristics and the stop decision (a stop, or the world's last step, frees the slot).
__global__ __launch_bounds__(kPairThreads) void run_stats_kernel(RunStatArgs a) {
  const int slot = blockIdx.x;
  const int32_t* list = a.par ? a.c.list1 : a.c.list0;
  if (slot >= a.c.nlive[a.par]) return;  // block-uniform
  const int64_t b = list[slot];
  const lx_sums s = lx_block_pair_sums(a.part + int64_t(slot) * a.C * a.pairs * LX_PAIR_REC, a.C, a.pairs);
  if (threadIdx.x != 0) return;
  lx_carry cr = a.c.carry[b];
  float* out = a.stats ? a.stats + (b * a.max_steps + cr.t) * (LX_NSTAT_BASE + a.C) : nullptr;
  const int stop = lx_stats_finish(s, cr, a.ndim, a.dims, a.C, a.R2, a.R, a.dt, a.stop_policy, a.min_steps,
                                   a.cells_total, out);
  const int done = stop || cr.t >= a.max_steps;
  a.c.carry[b] = cr;
  a.c.done[slot] = done;
  if (done) {
    if (a.n_out) a.n_out[b] = cr.n;
    if (a.steps_out) a.steps_out[b] = cr.t;
  }
}

// One block: the next step's live list, the slots whose world goes on (in order) and then worlds
// from the queue while slots are free.  NT threads: 1024, or one wave for the buckets of at most 64
// slots (the scan's rounds are then barriers of a single wave).
constexpr int kCtlThreads = 1024;
template <int NT>
__global__ __launch_bounds__(NT) void run_control_kernel(RunCtl c, int par, int slots, int64_t B) {
  __shared__ int scan[NT];
  __shared__ int64_t head_s;
  const int tid = threadIdx.x;
  const int32_t* in = par ? c.list1 : c.list0;
  int32_t* out = par ? c.list0 : c.list1;
  const int n = c.nlive[par];
  const int per = (n + NT - 1) / NT;
  const int s0 = min(tid * per, n), s1 = min(s0 + per, n);
  int cnt = 0;
  for (int s = s0; s < s1; ++s) cnt += c.done[s] ? 0 : 1;
  scan[tid] = cnt;
  if (tid == 0) head_s = *c.head;
  __syncthreads();
  for (int off = 1; off < NT; off <<= 1) {  // inclusive scan of the survivor counts
    const int v = tid >= off ? scan[tid - off] : 0;
    __syncthreads();
    scan[tid] += v;
    __syncthreads();
  }
  int o = scan[tid] - cnt;
  const int kept = scan[NT - 1];
  for (int s = s0; s < s1; ++s)
    if (!c.done[s]) out[o++] = in[s];
  const int64_t h = head_s;
  const int64_t room = slots - kept, left = B - h;
  const int take = int(room < left ? room : left);
  for (int i = tid; i < take; i += NT) out[kept + i] = int32_t(h + i);
  if (tid == 0) {
    *c.head = int32_t(h + take);
    c.nlive[par ^ 1] = kept + take;
  }
}

__global__ void zero_i32_kernel(int32_t* p, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) p[i] = 0;
}

// after the run: worlds that ran an odd number of steps hold their final state in X1
__global__ void run_finish_kernel(const lx_carry* __restrict__ carry, const float* __restrict__ X1,
                                  float* __restrict__ X0, int64_t B, int64_t elems) {
  const int64_t total = B * elems;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
    if (carry[i / elems].t & 1) X0[i] = X1[i];
}

struct PassesRun {
  const lx_rule* rule;
  const lx_run_args* a;
  Geom g;
  RunCtl ctl;
  RunView rv[2];       // per list parity
  float2* S;
  float2* U;
  float* part;
  int slots;
  int gsteps;  // steps per graph
};

// A line pass of the step over the live list, grids sized for `bucket` slots.  With `tune`, only the
// choice of its variant is made (timing the candidates, which write nothing but the slots' scratch
// spectra), outside any capture; otherwise the chosen variant is launched.
int run_lines(const PassesRun& r, const RunView& rv, int bucket, int axis, LineArgs la, int slot_imgs, bool tune,
              hipStream_t st) {
  const Geom& g = r.g;
  la.WH = g.WH; la.A = g.A[axis]; la.a_stride = g.a_stride[axis]; la.st = g.st[axis]; la.img_elems = g.half;
  la.k_elems = la.kquad ? int64_t(g.L[0] / 2 + 1) * g.WH : g.half;
  la.nlines = int64_t(slot_imgs) * g.A[axis] * g.WH;  // per slot: the kernels scale it by the live count
  la.list = rv.list; la.nlive = rv.nlive;
  const int L = g.L[axis];
  const int64_t lines = la.nlines * bucket;
  const LineVariant v = choose_lines(la, L, lines, true, tune, st,
                                     [&](LineVariant c) { return launch_lines_v<true>(la, L, lines, c, st); });
  return tune ? LX_OK : launch_lines_v<true>(la, L, lines, v, st);
}

// the variants of the step's line passes for `bucket` slots, chosen before a capture
int tune_run_lines(const PassesRun& r, int bucket, hipStream_t st) {
  const lx_rule* rule = r.rule;
  const int C = rule->C, K = rule->K;
  if (!lx_tune_should_time(st)) return LX_OK;
  int rc = LX_OK;
  if (r.g.naxes == 2) {
    LineArgs la = line_args();
    la.S = r.S; la.mode = 1;
    rc = run_lines(r, r.rv[0], bucket, 1, la, C, true, st);
  }
  if (!rc) {
    LineArgs la = line_args();
    la.S = r.S; la.U = r.U; la.Kh = static_cast<const float2*>(r.a->K_half_d); la.k_bstride = r.a->k_bstride;
    la.C = C; la.K = K; la.kquad = rule->k_quadrant;
    for (int k = 0; k < LX_MAX_K; ++k) la.cin[k] = k < K ? rule->cin[k] : -1;
    rc = run_lines(r, r.rv[0], bucket, 0, la, C, true, st);
  }
  if (!rc && r.g.naxes == 2) {
    LineArgs la = line_args();
    la.S = r.U; la.U = r.U; la.mode = 2;
    rc = run_lines(r, r.rv[0], bucket, 1, la, K, true, st);
  }
  return rc;
}

// one step over the live list of parity `par`, grids sized for `bucket` slots
int enqueue_step(const PassesRun& r, int par, int bucket, hipStream_t st) {
  const lx_rule* rule = r.rule;
  const lx_run_args* a = r.a;
  const Geom& g = r.g;
  const int C = rule->C, K = rule->K;
  const RunView& rv = r.rv[par];
  const int64_t slot_pairs = int64_t(C) * (g.rows / 2);
  int rc = dispatch_len(g.W, [&](auto w) {
    constexpr int W = decltype(w)::value;
    hipLaunchKernelGGL((rows_fwd_kernel<W, true>), dim3(blocks_for(bucket * slot_pairs * Lanes<W>::TPR)), dim3(256),
                       0, st, nullptr, r.S, slot_pairs, rv);
    return launched("rows_fwd (run)");
  });
  auto lines = [&](int axis, LineArgs la, int slot_imgs) {
    return run_lines(r, rv, bucket, axis, la, slot_imgs, false, st);
  };
  if (!rc && g.naxes == 2) {
    LineArgs la = line_args();
    la.S = r.S; la.mode = 1;
    rc = lines(1, la, C);
  }
  if (!rc) {
    LineArgs la = line_args();
    la.S = r.S; la.U = r.U; la.Kh = static_cast<const float2*>(a->K_half_d); la.k_bstride = a->k_bstride;
    la.C = C; la.K = K; la.kquad = rule->k_quadrant;
    for (int k = 0; k < LX_MAX_K; ++k) la.cin[k] = k < K ? rule->cin[k] : -1;
    rc = lines(0, la, C);
  }
  if (!rc && g.naxes == 2) {
    LineArgs la = line_args();
    la.S = r.U; la.U = r.U; la.mode = 2;
    rc = lines(1, la, K);
  }
  if (rc) return rc;
  RowArgs ra = row_args(rule);
  ra.U = r.U; ra.gf = a->gf_d; ra.gf_bs = a->gf_bstride; ra.w = a->w_d; ra.w_bs = a->w_bstride;
  ra.dt = a->dt_d; ra.dt_bs = a->dt_bstride;
  ra.update = 1; ra.part = r.part;
  ra.rv = rv;
  ra.field = a->field_hist_d; ra.pot = a->potential_hist_d; ra.cells_hist = a->cells_hist_d; ra.hist_B = a->B;
  ra.rows = g.rows;
  ra.npairs = slot_pairs;
  rc = dispatch_len(g.W, [&](auto w) {
    constexpr int W = decltype(w)::value;
    const dim3 grid(blocks_for(bucket * slot_pairs * Lanes<W>::TPR));
    const bool uni = (ra.rows / 2) % (64 / Lanes<W>::TPR) == 0;
    if (grid.x < 512) {  // the bucket of the last few worlds: latency-bound
      if (uni) hipLaunchKernelGGL((rows_inv_kernel<W, true, true, true, true>), grid, dim3(256), 0, st, ra);
      else hipLaunchKernelGGL((rows_inv_kernel<W, true, false, true, true>), grid, dim3(256), 0, st, ra);
    } else if (uni) {
      hipLaunchKernelGGL((rows_inv_kernel<W, true, true, true>), grid, dim3(256), 0, st, ra);
    } else {
      hipLaunchKernelGGL((rows_inv_kernel<W, true, false, true>), grid, dim3(256), 0, st, ra);
    }
    return launched("rows_inv (run)");
  });
  if (rc) return rc;
  RunStatArgs sa;
  sa.c = r.ctl; sa.par = par; sa.part = r.part; sa.pairs = g.rows / 2; sa.C = C; sa.ndim = rule->ndim;
  for (int i = 0; i < 3; ++i) sa.dims[i] = i < rule->ndim ? rule->dims[i] : 1;
  sa.R2 = (float)((double)a->R * (double)a->R); sa.R = a->R; sa.dt = (float)(1.0 / (double)a->T);
  sa.stop_policy = a->stop_policy; sa.min_steps = a->min_steps; sa.max_steps = a->max_steps;
  sa.cells_total = float(cells_per_world(rule) * C);
  sa.stats = a->stats_d; sa.n_out = a->n_d; sa.steps_out = a->steps_d;
  hipLaunchKernelGGL(run_stats_kernel, dim3(unsigned(bucket)), dim3(kPairThreads), 0, st, sa);
  if (bucket <= 64)
    hipLaunchKernelGGL(run_control_kernel<64>, dim3(1), dim3(64), 0, st, r.ctl, par, r.slots, a->B);
  else
    hipLaunchKernelGGL(run_control_kernel<kCtlThreads>, dim3(1), dim3(kCtlThreads), 0, st, r.ctl, par, r.slots, a->B);
  return launched("run statistics / control");
}

// twiddle tables of every length the step uses, built before a capture starts
int warm_tables(const Geom& g, hipStream_t st) {
  auto warm = [&](int L) { return dispatch_len(L, [&](auto l) { return ensure_tw<decltype(l)::value>(st); }); };
  if (int rc = warm(g.W)) return rc;
  for (int ax = 0; ax < g.naxes; ++ax)
    if (int rc = warm(g.L[ax])) return rc;
  return LX_OK;
}

hipStream_t capture_stream() {
  static hipStream_t s[kMaxDevices] = {};
  static std::mutex mu;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!s[dev] && hipStreamCreateWithFlags(&s[dev], hipStreamNonBlocking) != hipSuccess) s[dev] = nullptr;
  return s[dev];
}

// pinned host slots the live counts are copied into (two per device)
int32_t* live_count_slots() {
  static int32_t* p[kMaxDevices] = {};
  static std::mutex mu;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!p[dev] && hipHostMalloc(reinterpret_cast<void**>(&p[dev]), 2 * sizeof(int32_t), hipHostMallocDefault) !=
                     hipSuccess)
    p[dev] = nullptr;
  return p[dev];
}

struct GraphSet {
  std::vector<hipGraphExec_t> exec;  // by bucket index k: slots >> k
  ~GraphSet() {
    for (hipGraphExec_t e : exec)
      if (e) hipGraphExecDestroy(e);
  }
};

int capture(const PassesRun& r, int bucket, hipGraphExec_t* out, hipStream_t st) {
  if (int rc = tune_run_lines(r, bucket, st)) return rc;  // on the caller's stream, before the capture
  hipStream_t cs = capture_stream();
  if (!cs) return lx_fail(LX_EDEVICE, "lx_run: no capture stream");
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return lx_fail(LX_EDEVICE, "lx_run: hipStreamBeginCapture failed");
  int rc = LX_OK;
  for (int i = 0; i < r.gsteps && !rc; ++i) rc = enqueue_step(r, i & 1, bucket, cs);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(cs, &graph);
  if (rc) {
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess || !graph) return lx_fail(LX_EDEVICE, "lx_run: capture failed: %s", hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ei != hipSuccess) return lx_fail(LX_EDEVICE, "lx_run: hipGraphInstantiate: %s", hipGetErrorString(ei));
  return LX_OK;
}

}  // namespace

int lx_passes_scratch(const lx_rule* rule, int64_t B, int32_t slots, size_t* bytes) {
  Geom g;
  if (int rc = make_geom(rule->ndim, rule->dims, g)) return rc;
  const int64_t cap = slots > 0 && slots < B ? slots : B;
  *bytes = passes_layout(rule, g, B, cap < 1 ? 1 : cap).total;
  return LX_OK;
}

int lx_passes_run(const lx_rule* rule, const lx_run_args* a, hipStream_t st) {
  if (a->max_steps == 0) {  // no step: the initial states are the final ones
    const int64_t elems = int64_t(rule->C) * cells_per_world(rule);
    if (a->cells_d != a->cells0_d &&
        hipMemcpyAsync(a->cells_d, a->cells0_d, size_t(a->B * elems) * sizeof(float), hipMemcpyDeviceToDevice,
                       st) != hipSuccess)
      return lx_fail(LX_EDEVICE, "lx_run: copy of the initial states failed");
    if (a->n_d) hipLaunchKernelGGL(zero_i32_kernel, dim3(blocks_for(a->B)), dim3(256), 0, st, a->n_d, a->B);
    if (a->steps_d) hipLaunchKernelGGL(zero_i32_kernel, dim3(blocks_for(a->B)), dim3(256), 0, st, a->steps_d, a->B);
    return launched("run (no step)");
  }
  PassesRun r;
  r.rule = rule;
  r.a = a;
  if (int rc = make_geom(rule->ndim, rule->dims, r.g)) return rc;
  if (a->B > INT32_MAX) return lx_fail(LX_ENOTSUP, "lx_run: more than 2^31 - 1 worlds");
  const int64_t cap = a->slots > 0 && a->slots < a->B ? a->slots : a->B;
  if (cap > (1 << 20)) return lx_fail(LX_ENOTSUP, "lx_run: more than 2^20 worlds in flight (set slots)");
  r.slots = int(cap);
  r.gsteps = steps_per_graph(cap * rule->C * cells_per_world(rule));
  const PassesLayout l = passes_layout(rule, r.g, a->B, cap);
  char* scr = static_cast<char*>(a->scratch_d);
  r.ctl.list0 = reinterpret_cast<int32_t*>(scr + l.list0);
  r.ctl.list1 = reinterpret_cast<int32_t*>(scr + l.list1);
  r.ctl.nlive = reinterpret_cast<int32_t*>(scr + l.ctl);
  r.ctl.head = r.ctl.nlive + 2;
  r.ctl.done = reinterpret_cast<int32_t*>(scr + l.done);
  r.ctl.carry = reinterpret_cast<lx_carry*>(scr + l.carry);
  r.S = reinterpret_cast<float2*>(scr + l.S);
  r.U = reinterpret_cast<float2*>(scr + l.U);
  r.part = reinterpret_cast<float*>(scr + l.part);
  const int64_t elems = int64_t(rule->C) * cells_per_world(rule);
  float* X0 = a->cells_d;
  float* X1 = reinterpret_cast<float*>(scr + l.X1);
  for (int p = 0; p < 2; ++p)
    r.rv[p] = RunView{p ? r.ctl.list1 : r.ctl.list0, r.ctl.nlive + p, r.ctl.carry, X0, X1, elems};
  int rc = warm_tables(r.g, st);
  if (rc) return rc;
  if (a->cells_d != a->cells0_d &&
      hipMemcpyAsync(a->cells_d, a->cells0_d, size_t(a->B * elems) * sizeof(float), hipMemcpyDeviceToDevice, st) !=
          hipSuccess)
    return lx_fail(LX_EDEVICE, "lx_run: copy of the initial states failed");
  const unsigned sgrid = blocks_for(std::max<int64_t>(a->B, cap));
  hipLaunchKernelGGL(run_start_kernel, dim3(std::min(sgrid, 4096u)), dim3(256), 0, st, r.ctl, a->B, r.slots);
  if ((rc = launched("run start"))) return rc;

  int nb = 1;  // buckets: slots, slots / 2, ... down to 1
  while ((r.slots >> nb) >= 1 && nb < 31) ++nb;
  GraphSet gs;
  gs.exec.assign(nb, nullptr);
  int32_t* host_n = live_count_slots();
  if (!host_n) return lx_fail(LX_EDEVICE, "lx_run: pinned host memory unavailable");
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (int i = 0; i < 2; ++i)
    if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess)
      return lx_fail(LX_EDEVICE, "lx_run: event creation failed");
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() { for (int i = 0; i < 2; ++i) if (e[i]) hipEventDestroy(e[i]); }
  } evg{ev};
  // a bound on the replays: list scheduling finishes within (total work) / slots + the longest world
  const int64_t bound_steps = (a->B * int64_t(a->max_steps) + cap - 1) / cap + a->max_steps;
  const int64_t max_graphs = bound_steps / r.gsteps + 4;
  int64_t seen = cap;  // the live count the host last read (an upper bound of the next replays' counts)
  for (int64_t i = 0;; ++i) {
    if (i > max_graphs) return lx_fail(LX_EDEVICE, "lx_run: the live list did not drain (internal error)");
    int k = 0;
    while (k + 1 < nb && (int64_t(r.slots) >> (k + 1)) >= seen) ++k;
    if (!gs.exec[k] && (rc = capture(r, r.slots >> k, &gs.exec[k], st))) return rc;
    if (hipGraphLaunch(gs.exec[k], st) != hipSuccess) return lx_fail(LX_EDEVICE, "lx_run: hipGraphLaunch failed");
    // the live count after this replay (step parity 0 again: the graph has an even number of steps)
    if (hipMemcpyAsync(host_n + (i & 1), r.ctl.nlive, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipEventRecord(ev[i & 1], st) != hipSuccess)
      return lx_fail(LX_EDEVICE, "lx_run: live count copy failed");
    if (i == 0) continue;
    if (hipEventSynchronize(ev[(i - 1) & 1]) != hipSuccess) return lx_fail(LX_EDEVICE, "lx_run: stream failed");
    seen = host_n[(i - 1) & 1];
    if (seen == 0) break;
  }
  hipLaunchKernelGGL(run_finish_kernel, dim3(std::min(blocks_for(a->B * elems), 65536u)), dim3(256), 0, st,
                     r.ctl.carry, X1, X0, a->B, elems);
  return launched("run finish");
}

// ====================================================================== the reverse of one step
// lx_adjoint_step: lx_adjoint_pointwise then lx_adjoint_potential, with one channel and one kernel
// fused into a row pass that computes dL/dU from the potential and lambda (and the state where the
// update needs it) and transforms it at once: dL/dU is never written, the gradient scalars are
// summed per block, and lambda is rewritten in that pass only where the step changes it (v2, the
// true clip).  Several channels or kernels take the two entry points in turn.
namespace {

struct AdjRowArgs {
  const float* U;  // [B][rows][W] potentials
  float* lam;      // [B][rows][W] dL/d(state out) in, the direct part of dL/d(state in) out
  const float* A;  // states (read for v2 and the true clip)
  float2* S;       // [B][rows][W/2+1] half spectra of dL/dU along W
  float* part;     // [B][blocks][4] partial sums of dm, ds, dw, ddt
  const float* gf;
  int64_t gf_bs;
  const float* w;
  int64_t w_bs;
  const float* dt;
  int64_t dt_bs;
  int64_t rows;    // rows per world
  int state_fn, average, ste;
  int accumulate;  // 1: add the block sums into part (a run's steps), 0: overwrite
  // BRIDGE: the later step's adjoint potential, half spectra transformed along every axis but W
  // (may alias S); its inverse row transform is added to lam before the pointwise reverse
  const float2* Sin;
  float inv_n;     // 1 / cells per world
};

// grid (blocks per world, B): a block's row pairs lie in one world (pairs past the world's last are
// idle lanes: they load pair 0, contribute nothing and store nothing).
// BRIDGE: the last row pass of one reverse step (lambda += inverse transform of its adjoint
// potential, as rows_inv_kernel in selection mode) and the first of the step before it in one
// kernel: lambda of the later step never leaves the registers between the two.  Every row pair is
// read and written by its own lane group only, so Sin may alias S.
template <int W, int KIND, bool BRIDGE = false>
__global__ __launch_bounds__(256) void rows_fwd_adj_kernel(AdjRowArgs a) {
  using LC = Lanes<W>;
  constexpr int TPR = LC::TPR, P = LC::P, WH = W / 2 + 1;
  __shared__ __align__(16) float2 xch_all[LC::XCH];
  const int lane = threadIdx.x & 63;
  const int t = threadIdx.x % TPR;
  float2* xch = xch_all + (threadIdx.x / TPR) * (TPR * LC::XRS);
  const int64_t b = blockIdx.y;
  const int64_t pr = int64_t(blockIdx.x) * (256 / TPR) + threadIdx.x / TPR;  // row pair in the world
  const bool live = pr < a.rows / 2;
  const int64_t row0 = live ? pr * 2 : 0;
  const int64_t off = (b * a.rows + row0) * W;
  const float w = a.w[b * a.w_bs];
  const float gm = a.gf[b * a.gf_bs], gs = a.gf[b * a.gf_bs + 1];
  const float inv_s = 1.f / gs, inv_9s2 = 1.f / (9.f * (gs * gs));
  const float dtw = a.dt[b * a.dt_bs];
  const float coef = a.average ? w / w : w;
  const float inv_den = a.average ? 1.f / w : 1.f;
  // mode 0: lambda passes through (straight-through v1, simple); 1: the true clip; 2: v2
  const int mode = a.state_fn == LX_STATE_V2 ? 2 : (a.state_fn == LX_STATE_V1 && !a.ste ? 1 : 0);
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  float2 x[P];
  // one base pointer per array: every point is a constant offset from it
  const float* pU = a.U + off + t;
  float* pL = a.lam + off + t;
  const float* pA = a.A + off + t;
  // every point's potential and lambda loaded at once (their latency overlaps the bridge's inverse
  // transform, and the loop below waits once instead of once per point)
  float uv[2][P], lv[2][P];
#pragma unroll
  for (int m = 0; m < P; ++m) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      uv[r][m] = pU[r * W + TPR * m];
      lv[r][m] = pL[r * W + TPR * m];
    }
  }
  if constexpr (BRIDGE) {
    // the later step's adjoint potential along W: stage the pair's half spectra, repack them as one
    // complex row Z[k] = A[k] + i B[k] (A[k] = conj(A[W-k]) above W/2), inverse transform
    const float2* u0 = a.Sin + (b * a.rows + row0) * int64_t(WH);
    const float2* u1 = u0 + WH;
#pragma unroll
    for (int i = 0; i < (WH + TPR - 1) / TPR; ++i) {
      const int col = t + TPR * i;
      if (col < WH) {
        xch[col] = u0[col];
        xch[WH + col] = u1[col];
      }
    }
    wave_lds_fence();
#pragma unroll
    for (int e = 0; e < P; ++e) {
      const int kk = freq_of<W>(t, e);
      const bool mirror = kk >= W / 2;
      const float sg = mirror ? -1.f : 1.f;
      const int col = mirror ? W - kk : kk;
      const float2 ya = xch[col], yb = xch[WH + col];
      float2 z = make_float2(fmaf(-sg, yb.y, ya.x), fmaf(sg, ya.y, yb.x));
      if (kk == 0 || kk == W / 2) z = make_float2(ya.x, yb.x);
      x[e] = z;
    }
    wave_lds_fence();
    fft_group<W, true>(x, t, lane, xch);
    // parked in the group's exchange rows (point TPR m + t of both rows at TPR m + t) until the
    // pointwise loop adds them to lambda: the registers are free for that loop
    wave_lds_fence();
#pragma unroll
    for (int m = 0; m < P; ++m) xch[TPR * m + t] = x[m];
    wave_lds_fence();
    __builtin_amdgcn_sched_barrier(0);
  }
  // the pointwise reverse of every point the lanes hold, straight-line for each update rule (the
  // rule and the weighting are uniform: branches around the loop, not in it); the field is the
  // kernel's growth recomputed from the potential as the forward computes it
  auto points = [&](auto mode_c, auto avg_c) {
    constexpr int MODE = decltype(mode_c)::value;
    constexpr bool AVG = decltype(avg_c)::value != 0;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      float du[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int o = r * W + TPR * m;
        const float U = uv[r][m];
        float L = lv[r][m];
        if constexpr (BRIDGE) {
          // dL/d(state in) of the later step: its direct part plus the correlation term, as
          // rows_inv_kernel's selection mode adds them
          const float2 inc = xch[TPR * m + t];
          L = add_nc(L, mul_nc(r == 0 ? inc.x : inc.y, a.inv_n));
          if (MODE == 0 && live) pL[o] = L;
        }
        const float A = MODE != 0 ? pA[o] : 0.f;
        float dgdu, dgdm, dgds;
        const float g = growth_rcp(KIND, gm, gs, inv_s, inv_9s2, U);
        growth_grad(KIND, gm, gs, inv_s, inv_9s2, U, dgdu, dgdm, dgds);
        const float wg = mul_nc(w, g);
        const float G = AVG ? wg / w : wg;
        float dz = live ? L : 0.f;
        if constexpr (MODE == 1) {
          const float z = add_nc(A, mul_nc(dtw, G));
          dz = (live && z > 0.f && z < 1.f) ? L : 0.f;
        }
        const float dG = dtw * dz;
        const float dg = dG * coef;
        du[r] = dg * dgdu;
        acc[0] += dg * dgdm;
        acc[1] += dg * dgds;
        acc[2] += dG * (AVG ? (g - G) * inv_den : g);
        if constexpr (MODE == 2) {
          if (live) pL[o] = dz * (1.f - dtw);
          acc[3] += dz * (G - A);
        } else {
          if constexpr (MODE == 1) {
            if (live) pL[o] = dz;
          }
          acc[3] += dz * G;
        }
      }
      // the point's arithmetic is done here, next to its loads: left free, the compiler sinks it
      // to the transform below and keeps every point's loaded values live until then
      __asm__ volatile("" : "+v"(du[0]), "+v"(du[1]), "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]));
      x[m] = make_float2(du[0], du[1]);
      if ((m & 1) == 1) __builtin_amdgcn_sched_barrier(0);
    }
  };
  if (mode == 0) {
    if (a.average) points(IntC<0>{}, IntC<1>{}); else points(IntC<0>{}, IntC<0>{});
  } else if (mode == 1) {
    if (a.average) points(IntC<1>{}, IntC<1>{}); else points(IntC<1>{}, IntC<0>{});
  } else {
    if (a.average) points(IntC<2>{}, IntC<1>{}); else points(IntC<2>{}, IntC<0>{});
  }
  fft_group<W, false>(x, t, lane, xch);
  wave_lds_fence();
#pragma unroll
  for (int e = 0; e < P; ++e) xch[freq_of<W>(t, e)] = x[e];
  wave_lds_fence();
  float2* s0 = a.S + (b * a.rows + row0) * int64_t(WH);
  float2* s1 = s0 + WH;
#pragma unroll
  for (int i = 0; i < (WH + TPR - 1) / TPR; ++i) {
    const int k = t + TPR * i;
    if (k < WH && live) {
      const float2 z = xch[k];
      const float2 zn = xch[k == 0 ? 0 : W - k];
      s0[k] = make_float2(0.5f * (z.x + zn.x), 0.5f * (z.y - zn.y));
      s1[k] = make_float2(0.5f * (z.y + zn.y), 0.5f * (zn.x - z.x));
    }
  }
  lx_block_partial<4>(acc, a.part, a.accumulate != 0);
}

}  // namespace

int lx_adj11_init(const lx_rule* rule, int64_t B, const float* gf_d, int64_t gf_bstride, const float* w_d,
                  int64_t w_bstride, const float* dt_d, int64_t dt_bstride, const void* K_half_d, int64_t k_bstride,
                  int32_t ste, float* lambda_d, float* grad_gf_d, float* grad_w_d, float* grad_dt_d, void* work_d,
                  void* adj_work_d, void* stream, LxAdj11* c) {
  Geom g;
  if (int rc = make_geom(rule->ndim, rule->dims, g)) return rc;
  if (!(rule->C == 1 && rule->K == 1) || B < 1 || B > 65535)
    return lx_fail(LX_EINVAL, "fused reverse step: one channel, one kernel, 1 to 65535 worlds");
  c->rule = rule; c->B = B; c->stream = stream;
  c->gf = gf_d; c->gf_bs = gf_bstride; c->w = w_d; c->w_bs = w_bstride; c->dt = dt_d; c->dt_bs = dt_bstride;
  c->K_half = K_half_d; c->k_bs = k_bstride; c->ste = ste;
  c->lam = lambda_d; c->g_gf = grad_gf_d; c->g_w = grad_w_d; c->g_dt = grad_dt_d;
  c->X = work_d;
  c->part = lx_adj_work_11(rule, B, adj_work_d);
  const int64_t pairs_per_block = 256 / lx_row_tpr(g.W);
  c->bpw = int((g.rows / 2 + pairs_per_block - 1) / pairs_per_block);
  return LX_OK;
}

// rows_fwd_adj_kernel over the step whose potential and input state are U and A (bridge: after
// the inverse row transform of the later step's adjoint potential in X)
int lx_adj11_rows(const LxAdj11& c, const float* U, const float* A, bool bridge, bool accumulate) {
  Geom g;
  if (int rc = make_geom(c.rule->ndim, c.rule->dims, g)) return rc;
  hipStream_t st = (hipStream_t)c.stream;
  float2* X = static_cast<float2*>(c.X);
  AdjRowArgs ar;
  ar.U = U; ar.lam = c.lam; ar.A = A; ar.S = X; ar.part = c.part;
  ar.gf = c.gf; ar.gf_bs = c.gf_bs; ar.w = c.w; ar.w_bs = c.w_bs; ar.dt = c.dt; ar.dt_bs = c.dt_bs;
  ar.rows = g.rows; ar.state_fn = c.rule->state_fn; ar.average = c.rule->average; ar.ste = c.ste != 0;
  ar.accumulate = accumulate ? 1 : 0;
  ar.Sin = X;
  ar.inv_n = 1.f / float(g.rows * g.W);
  return dispatch_len(g.W, [&](auto wc) {
    constexpr int W = decltype(wc)::value;
    if (int r2 = ensure_tw<W>(st)) return r2;
    const dim3 grid(unsigned(c.bpw), unsigned(c.B));
    auto go = [&](auto kind_c) {
      constexpr int KIND = decltype(kind_c)::value;
      if (bridge) hipLaunchKernelGGL((rows_fwd_adj_kernel<W, KIND, true>), grid, dim3(256), 0, st, ar);
      else hipLaunchKernelGGL((rows_fwd_adj_kernel<W, KIND, false>), grid, dim3(256), 0, st, ar);
    };
    switch (c.rule->gf[0]) {
      case LX_GF_POLY_QUAD4: go(IntC<LX_GF_POLY_QUAD4>{}); break;
      case LX_GF_GAUSSIAN: go(IntC<LX_GF_GAUSSIAN>{}); break;
      case LX_GF_GAUSSIAN_TARGET: go(IntC<LX_GF_GAUSSIAN_TARGET>{}); break;
      case LX_GF_STEP: go(IntC<LX_GF_STEP>{}); break;
      case LX_GF_STAIRCASE: go(IntC<LX_GF_STAIRCASE>{}); break;
      case LX_GF_TRIANGLE: go(IntC<LX_GF_TRIANGLE>{}); break;
      default: go(IntC<LX_GF_IDENTITY>{}); break;
    }
    return launched(bridge ? "rows (adjoint bridge)" : "rows_fwd (adjoint)");
  });
}

// the adjoint potential's passes between the row passes: lines, conj(K) (and the kernel gradient), lines
int lx_adj11_lines(const LxAdj11& c, const void* state_spec_d, void* dspec_d) {
  Geom g;
  int rc = make_geom(c.rule->ndim, c.rule->dims, g);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)c.stream;
  float2* X = static_cast<float2*>(c.X);
  if (g.naxes == 2) {
    LineArgs la = line_args();
    la.S = X; la.mode = 1;
    if ((rc = launch_lines(g, 1, la, c.B, st))) return rc;
  }
  {
    LineArgs la = line_args();
    la.S = X; la.U = X;
    la.Kh = static_cast<const float2*>(c.K_half); la.k_bstride = c.k_bs; la.C = 1; la.K = 1; la.conjk = 1;
    la.cin[0] = 0; la.acc_cin[0] = 0;
    la.FAin = static_cast<const float2*>(state_spec_d);
    la.dspec = static_cast<float2*>(dspec_d);
    la.C_fa = 1;
    if ((rc = launch_lines(g, 0, la, c.B, st))) return rc;
  }
  if (g.naxes == 2) {
    LineArgs la = line_args();
    la.S = X; la.U = X; la.mode = 2;
    if ((rc = launch_lines(g, 1, la, c.B, st))) return rc;
  }
  return LX_OK;
}

// the last row pass of the earliest step: lambda += the inverse row transform of X; finish: the
// gradient scalars summed from the partial sums (by extra workgroups of the same launch)
int lx_adj11_last(const LxAdj11& c, bool finish) {
  Geom g;
  if (int rc = make_geom(c.rule->ndim, c.rule->dims, g)) return rc;
  RowArgs ra = row_args(c.rule);
  ra.U = static_cast<float2*>(c.X); ra.A = c.lam; ra.A_out = c.lam;
  ra.update = 1; ra.sel = 1; ra.average = 0; ra.state_fn = LX_STATE_SIMPLE;
  if (finish) {
    LxAdjFinish& f = ra.fin;
    f.part = c.part; f.B = c.B; f.per = c.bpw;
    f.g_gf = c.g_gf; f.gf_bs = c.gf_bs; f.g_w = c.g_w; f.w_bs = c.w_bs; f.g_dt = c.g_dt; f.dt_bs = c.dt_bs;
    const int64_t nout = (c.g_gf ? (c.gf_bs ? c.B : 1) * 2 : 0) + (c.g_w ? (c.w_bs ? c.B : 1) : 0) +
                         (c.g_dt ? (c.dt_bs ? c.B : 1) : 0);
    f.blocks = int((nout + 3) / 4);
  }
  return launch_rows_inv(g, ra, c.B, (hipStream_t)c.stream);
}

extern "C" int lx_adjoint_step(const lx_rule* rule, int64_t B, const float* state_d, const float* field_d,
                               const float* potential_d, const float* gf_d, int64_t gf_bstride, const float* w_d,
                               int64_t w_bstride, const float* dt_d, int64_t dt_bstride, const void* K_half_d,
                               int64_t k_bstride, int32_t ste, const void* state_spec_d, void* dspec_d,
                               float* lambda_d, float* dpot_d, float* grad_gf_d, float* grad_w_d, float* grad_dt_d,
                               void* work_d, void* adj_work_d, void* stream) {
  int rc = lx_check_rule(rule);
  if (rc || (rc = lx_no_quadrant(rule, "lx_adjoint_step"))) return rc;
  if ((state_spec_d == nullptr) != (dspec_d == nullptr))
    return lx_fail(LX_EINVAL, "state_spec and dspec are given together or not at all");
  Geom g;
  if ((rc = make_geom(rule->ndim, rule->dims, g))) return rc;
  if (B < 0) return lx_fail(LX_EINVAL, "B < 0");
  if (B == 0) return LX_OK;
  const bool one = rule->C == 1 && rule->K == 1;
  if (!state_d || !potential_d || !gf_d || !w_d || !dt_d || !K_half_d || !lambda_d || !work_d || !adj_work_d ||
      (!one && (!field_d || !dpot_d)))
    return lx_fail(LX_EINVAL, "NULL pointer");
  const bool fused = one && B <= 65535;
  if (!fused) {
    if (!dpot_d) return lx_fail(LX_EINVAL, "lx_adjoint_step: this rule needs dpot_d");
    if ((rc = lx_adjoint_pointwise(rule, B, state_d, field_d, potential_d, gf_d, gf_bstride, w_d, w_bstride, dt_d,
                                   dt_bstride, lambda_d, nullptr, nullptr, ste, lambda_d, dpot_d, grad_gf_d, grad_w_d,
                                   grad_dt_d, adj_work_d, stream)))
      return rc;
    return lx_adjoint_potential(rule, B, K_half_d, k_bstride, dpot_d, state_spec_d, dspec_d, lambda_d, work_d,
                                stream);
  }
  LxAdj11 c;
  if ((rc = lx_adj11_init(rule, B, gf_d, gf_bstride, w_d, w_bstride, dt_d, dt_bstride, K_half_d, k_bstride, ste,
                          lambda_d, grad_gf_d, grad_w_d, grad_dt_d, work_d, adj_work_d, stream, &c)))
    return rc;
  if ((rc = lx_adj11_rows(c, potential_d, state_d, false, false))) return rc;
  if ((rc = lx_adj11_lines(c, state_spec_d, dspec_d))) return rc;
  return lx_adj11_last(c, true);
}
