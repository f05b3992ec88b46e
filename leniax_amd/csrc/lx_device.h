// Device-side building blocks for the Lenia FFT step on CDNA4 (gfx950).
//
//  * complex helpers on float2
//  * compile-time twiddle constants (constexpr double series, rounded to float)
//  * fft_reg<P>:       in-register radix-2 FFT of P points, natural order in/out
//  * fft_dist<P,INV>:  a length-4P FFT spread over the 4 lanes of a DPP quad:
//                      in-register P-point FFTs + one radix-4 step whose data
//                      exchange uses DPP quad permutes (no LDS traffic)
//  * growth functions (leniax/growth_functions.py) by id
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LX_DEV __device__ __forceinline__

// ----------------------------------------------------------------- complex
LX_DEV float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
LX_DEV float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
LX_DEV float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
LX_DEV float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
LX_DEV float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
LX_DEV float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// ----------------------------------------------------------------- constexpr twiddles
// sin/cos of 2*pi*k/n from the Taylor series on [0, pi/4] after exact octant
// reduction in integers; double precision, then rounded to float.
namespace lxc {
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double sin_series(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int n = 1; n < 14; ++n) { term *= -x2 / double((2 * n) * (2 * n + 1)); sum += term; }
  return sum;
}
constexpr double cos_series(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int n = 1; n < 14; ++n) { term *= -x2 / double((2 * n - 1) * (2 * n)); sum += term; }
  return sum;
}
// cos(2*pi*k/n), sin(2*pi*k/n) with k reduced into the first octant.
constexpr double cospi2(long k, long n) {
  k %= n; if (k < 0) k += n;
  // angle = 2*pi*k/n; work with 8k vs n to find the octant
  long k8 = 8 * k;
  long oct = k8 / n;           // 0..7
  // reduce by symmetries: cos(a) for a in octant
  switch (oct) {
    case 0: return cos_series(2 * kPi * double(k) / double(n));
    case 1: return sin_series(2 * kPi * double(n - 4 * k) / double(4 * n));           // cos(a)=sin(pi/2-a)
    case 2: return -sin_series(2 * kPi * double(4 * k - n) / double(4 * n));          // a-pi/2
    case 3: return -cos_series(2 * kPi * double(n - 2 * k) / double(2 * n));          // pi-a
    case 4: return -cos_series(2 * kPi * double(2 * k - n) / double(2 * n));          // a-pi
    case 5: return -sin_series(2 * kPi * double(3 * n - 4 * k) / double(4 * n));      // 3pi/2-a
    case 6: return sin_series(2 * kPi * double(4 * k - 3 * n) / double(4 * n));       // a-3pi/2
    default: return cos_series(2 * kPi * double(n - k) / double(n));                  // 2pi-a
  }
}
}  // namespace lxc

// Forward twiddle W_n^k = exp(-2*pi*i*k/n); INV flips the sign of the angle.
template <int N, int K, bool INV>
struct Tw {
  static constexpr float re = float(lxc::cospi2(K, N));
  // sin(2*pi*k/n) = cos(2*pi*(4k - n)/(4n))
  static constexpr float s = float(lxc::cospi2(long(K) * 4 - N, 4L * N));
  static constexpr float im = INV ? s : -s;
};

// ----------------------------------------------------------------- in-register FFT
constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }
constexpr int bitrev_c(int i, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
  return r;
}

template <int LEN, int K, bool INV>
LX_DEV void bfly(float2& a, float2& b) {
  // a' = a + w b, b' = a - w b with w = W_LEN^K (compile-time constant)
  float2 wb;
  if constexpr (K == 0) {
    wb = b;
  } else if constexpr (4 * K == LEN) {  // w = -i (fwd) / +i (inv)
    wb = INV ? make_float2(-b.y, b.x) : make_float2(b.y, -b.x);
  } else {
    constexpr float wr = Tw<LEN, K, INV>::re;
    constexpr float wi = Tw<LEN, K, INV>::im;
    wb = make_float2(fmaf(b.x, wr, -b.y * wi), fmaf(b.x, wi, b.y * wr));
  }
  float2 t = a;
  a = cadd(t, wb);
  b = csub(t, wb);
}

template <int P, bool INV, int LEN, int I, int K>
LX_DEV void fft_stage_body(float2* x) {
  if constexpr (I < P) {
    if constexpr (K < LEN / 2) {
      bfly<LEN, K, INV>(x[I + K], x[I + K + LEN / 2]);
      fft_stage_body<P, INV, LEN, I, K + 1>(x);
    } else {
      fft_stage_body<P, INV, LEN, I + LEN, 0>(x);
    }
  }
}

template <int P, bool INV, int LEN>
LX_DEV void fft_stages(float2* x) {
  if constexpr (LEN <= P) {
    fft_stage_body<P, INV, LEN, 0, 0>(x);
    fft_stages<P, INV, LEN * 2>(x);
  }
}

template <int P, int I>
LX_DEV void bitrev_perm(float2* x) {
  if constexpr (I < P) {
    constexpr int J = bitrev_c(I, ilog2c(P));
    if constexpr (J > I) { float2 t = x[I]; x[I] = x[J]; x[J] = t; }
    bitrev_perm<P, I + 1>(x);
  }
}

constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

template <int P, bool INV>
LX_DEV void fft_reg(float2* x);

// Radix-3 step of a length-P FFT (P = 3Q): X[k1 + Q k2] = sum_r W_3^(r k2) W_P^(r k1) Y_r[k1],
// Y_r the length-Q transform of x[3 n + r].  Twiddles are compile-time constants.
template <int P, bool INV, int K1>
LX_DEV void radix3_body(const float2* y0, const float2* y1, const float2* y2, float2* x) {
  constexpr int Q = P / 3;
  if constexpr (K1 < Q) {
    const float2 a = y0[K1];
    float2 b = y1[K1], c = y2[K1];
    if constexpr (K1 != 0) {
      b = cmul(b, make_float2(Tw<P, K1, INV>::re, Tw<P, K1, INV>::im));
      c = cmul(c, make_float2(Tw<P, 2 * K1, INV>::re, Tw<P, 2 * K1, INV>::im));
    }
    constexpr float r = 0.866025403784438646763723170752936183f;  // sqrt(3) / 2
    const float2 s = cadd(b, c), d = csub(b, c);
    const float2 m = make_float2(fmaf(-0.5f, s.x, a.x), fmaf(-0.5f, s.y, a.y));
    // forward W_3 = -1/2 - i r: X1 = m - i r d, X2 = m + i r d (inverse: conjugate)
    const float2 ird = INV ? make_float2(-r * d.y, r * d.x) : make_float2(r * d.y, -r * d.x);
    x[K1] = cadd(a, s);
    x[K1 + Q] = cadd(m, ird);
    x[K1 + 2 * Q] = csub(m, ird);
    radix3_body<P, INV, K1 + 1>(y0, y1, y2, x);
  }
}

// Natural-order in, natural-order out, unnormalised; INV uses exp(+2*pi*i*jk/P).
// P a power of two (radix 2) or three times one (one radix-3 step on top).
template <int P, bool INV>
LX_DEV void fft_reg(float2* x) {
  if constexpr (is_pow2(P)) {
    bitrev_perm<P, 0>(x);
    fft_stages<P, INV, 2>(x);
  } else {
    static_assert(P % 3 == 0 && is_pow2(P / 3), "in-register FFT lengths are 2^k and 3 * 2^k");
    constexpr int Q = P / 3;
    float2 y0[Q], y1[Q], y2[Q];
#pragma unroll
    for (int n = 0; n < Q; ++n) {
      y0[n] = x[3 * n];
      y1[n] = x[3 * n + 1];
      y2[n] = x[3 * n + 2];
    }
    fft_reg<Q, INV>(y0);
    fft_reg<Q, INV>(y1);
    fft_reg<Q, INV>(y2);
    radix3_body<P, INV, 0>(y0, y1, y2, x);
  }
}

// ----------------------------------------------------------------- DPP quad exchange
// quad_perm(a,b,c,d): lane i of each quad reads lane perm[i] of the same quad.
#define LX_QP(a, b, c, d) ((a) | ((b) << 2) | ((c) << 4) | ((d) << 6))

template <int CTRL>
LX_DEV float dpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
LX_DEV float2 dpp2(float2 v) { return make_float2(dpp<CTRL>(v.x), dpp<CTRL>(v.y)); }

// Length-4P FFT over the 4 lanes of a quad (q = lane & 3), unnormalised.
//  forward: in  x[m]  = data[4m + q]                 (m = 0..P-1)
//           out x[k1] = DATA[k1 + P*bitrev2(q)]     (k1 = 0..P-1)
//  inverse: in  x[k1] = DATA[k1 + P*bitrev2(q)]
//           out x[m]  = data[4m + q]
// ``tw`` points at exp(-2*pi*i*j/(4P)), j = 0..4P-1 (LDS table).
template <int P, bool INV>
LX_DEV void fft_dist(float2* x, int q, const float2* tw) {
  constexpr int N = 4 * P;
  const float sgn_hi = (q & 2) ? -1.f : 1.f;  // (-1)^(q>>1)
  const float sgn_lo = (q & 1) ? -1.f : 1.f;  // (-1)^(q&1)
  if constexpr (!INV) {
    fft_reg<P, false>(x);
#pragma unroll
    for (int k = 1; k < P; ++k) x[k] = cmul(x[k], tw[(q * k) & (N - 1)]);
    // stage A: v_{t0}[s0] = y_{t0} + (-1)^{s0} y_{t0+2}, held at lane t0 + 2 s0
#pragma unroll
    for (int k = 0; k < P; ++k) {
      float2 a = dpp2<LX_QP(0, 1, 0, 1)>(x[k]);
      float2 b = dpp2<LX_QP(2, 3, 2, 3)>(x[k]);
      x[k] = make_float2(fmaf(sgn_hi, b.x, a.x), fmaf(sgn_hi, b.y, a.y));
    }
    // stage B: X_{s0+2s1} = v_0[s0] + W4^{s0} (-1)^{s1} v_1[s0], held at lane 2 s0 + s1
    // coefficient c: q=0 -> 1, q=1 -> -1, q=2 -> -i, q=3 -> +i
    const float cr = (q & 2) ? 0.f : sgn_lo;
    const float ci = (q & 2) ? -sgn_lo : 0.f;
#pragma unroll
    for (int k = 0; k < P; ++k) {
      float2 a = dpp2<LX_QP(0, 0, 2, 2)>(x[k]);
      float2 b = dpp2<LX_QP(1, 1, 3, 3)>(x[k]);
      x[k] = make_float2(fmaf(-ci, b.y, fmaf(cr, b.x, a.x)), fmaf(ci, b.x, fmaf(cr, b.y, a.y)));
    }
  } else {
    // inner: u_{s0}[t0] = Q_{s0} + (-1)^{t0} Q_{s0+2}; Q_s sits at lane bitrev2(s)
#pragma unroll
    for (int k = 0; k < P; ++k) {
      float2 a = dpp2<LX_QP(0, 0, 2, 2)>(x[k]);
      float2 b = dpp2<LX_QP(1, 1, 3, 3)>(x[k]);
      x[k] = make_float2(fmaf(sgn_lo, b.x, a.x), fmaf(sgn_lo, b.y, a.y));
    }
    // outer: z_{t0+2t1} = u_0[t0] + W4^{-t0} (-1)^{t1} u_1[t0], output at lane t
    // coefficient c: q=0 -> 1, q=1 -> +i, q=2 -> -1, q=3 -> -i
    const float cr = (q & 1) ? 0.f : sgn_hi;
    const float ci = (q & 1) ? sgn_hi : 0.f;
#pragma unroll
    for (int k = 0; k < P; ++k) {
      float2 a = dpp2<LX_QP(0, 1, 0, 1)>(x[k]);
      float2 b = dpp2<LX_QP(2, 3, 2, 3)>(x[k]);
      x[k] = make_float2(fmaf(-ci, b.y, fmaf(cr, b.x, a.x)), fmaf(ci, b.x, fmaf(cr, b.y, a.y)));
    }
#pragma unroll
    for (int k = 1; k < P; ++k) x[k] = cmulc(x[k], tw[(q * k) & (N - 1)]);
    fft_reg<P, true>(x);
  }
}

// ----------------------------------------------------------------- generic lane-group FFT
// Partner value of lane t ^ H inside groups of 2H lanes (H = 1, 2: DPP quad permutes;
// H = 4: row shifts by 4; H = 8: row rotate by 8).
template <int H>
LX_DEV float xor_fetch(float v, int lane) {
  if constexpr (H == 1) return dpp<LX_QP(1, 0, 3, 2)>(v);
  else if constexpr (H == 2) return dpp<LX_QP(2, 3, 0, 1)>(v);
  else if constexpr (H == 4) {
    const float up = dpp<0x104>(v);   // row_shl:4 -> lane i reads i+4
    const float dn = dpp<0x114>(v);   // row_shr:4 -> lane i reads i-4
    return (lane & 4) ? dn : up;
  } else {
    static_assert(H == 8, "lane groups up to 16");
    return dpp<0x128>(v);             // row_ror:8 -> lane i reads (i+8) mod 16
  }
}
template <int H>
LX_DEV float2 xor_fetch2(float2 v, int lane) { return make_float2(xor_fetch<H>(v.x, lane), xor_fetch<H>(v.y, lane)); }

constexpr int bitrev_n(int v, int n) {  // reverse the low log2(n) bits
  int r = 0;
  for (int b = 1; b < n; b <<= 1) { r <<= 1; r |= (v & 1); v >>= 1; }
  return r;
}

// Cross-lane DFT of size TPR over the lanes of a group, as a radix-2 network.
// forward (decimation in frequency): natural lane order in, X_s at lane bitrev(s) out.
// inverse (decimation in time):      X_s at lane bitrev(s) in, natural order out, x TPR.
// Stage H pairs lane t with t ^ H; the upper lane's twiddle W_{2H}^(t mod H) is read from
// the length-L table ``tw``.  Lower lanes use the twiddle 1 (an exact multiply), so the
// stage needs no lane selects.  SPREAD = true: for 8-lane groups whose bit 2 is lane
// bit 3 (lanes {0..3, 8..11} + offset), so that t ^ 4 is one DPP row rotate.
template <int H, bool SPREAD>
LX_DEV float2 group_xor(float2 v, int lane) {
  if constexpr (SPREAD && H == 4) return dpp2<0x128>(v);  // row_ror:8 == lane ^ 8
  else return xor_fetch2<H>(v, lane);
}

template <int P, int TPR, int L, bool INV, int H, bool SPREAD>
LX_DEV void lane_stage(float2* x, int t, int lane, const float2* tw) {
  const bool upper = (t & H) != 0;
  const float sgn = upper ? -1.f : 1.f;
  float2 w = make_float2(1.f, 0.f);
  if constexpr (H > 1) {
    const float2 tv = tw[(t & (H - 1)) * (L / (2 * H))];
    w = upper ? make_float2(tv.x, INV ? -tv.y : tv.y) : w;
  }
  if constexpr (!INV) {
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const float2 p = group_xor<H, SPREAD>(x[k], lane);
      const float2 v = make_float2(fmaf(sgn, x[k].x, p.x), fmaf(sgn, x[k].y, p.y));
      x[k] = (H > 1) ? cmul(v, w) : v;
    }
  } else {
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const float2 o = (H > 1) ? cmul(x[k], w) : x[k];
      const float2 p = group_xor<H, SPREAD>(o, lane);
      x[k] = make_float2(fmaf(sgn, o.x, p.x), fmaf(sgn, o.y, p.y));
    }
  }
}

template <int P, int TPR, int L, int H, bool SPREAD>
LX_DEV void lane_network_fwd(float2* x, int t, int lane, const float2* tw) {
  if constexpr (H >= 1) {
    lane_stage<P, TPR, L, false, H, SPREAD>(x, t, lane, tw);
    lane_network_fwd<P, TPR, L, H / 2, SPREAD>(x, t, lane, tw);
  }
}

template <int P, int TPR, int L, int H, bool SPREAD>
LX_DEV void lane_network_inv(float2* x, int t, int lane, const float2* tw) {
  if constexpr (H < TPR) {
    lane_stage<P, TPR, L, true, H, SPREAD>(x, t, lane, tw);
    lane_network_inv<P, TPR, L, H * 2, SPREAD>(x, t, lane, tw);
  }
}

// Length-(TPR*P) FFT over a group of TPR lanes, unnormalised.
//  forward: in  x[m]  = data[TPR*m + t];  out x[k1] = DATA[k1 + P*bitrev(t)]
//  inverse: in  x[k1] = DATA[k1 + P*bitrev(t)];  out x[m] = data[TPR*m + t]
// ``tw`` = exp(-2*pi*i*j/(TPR*P)), j < TPR*P (LDS or global table).
template <int P, int TPR, bool INV, bool SPREAD = false>
LX_DEV void fft_lanes(float2* x, int t, int lane, const float2* tw) {
  constexpr int L = TPR * P;
  if constexpr (!INV) {
    fft_reg<P, false>(x);
    int j = 0;
#pragma unroll
    for (int k = 1; k < P; ++k) {
      j += t;  // t * k < L
      x[k] = cmul(x[k], tw[j]);
    }
    lane_network_fwd<P, TPR, L, TPR / 2, SPREAD>(x, t, lane, tw);
  } else {
    lane_network_inv<P, TPR, L, 1, SPREAD>(x, t, lane, tw);
    int j = 0;
#pragma unroll
    for (int k = 1; k < P; ++k) {
      j += t;
      x[k] = cmulc(x[k], tw[j]);
    }
    fft_reg<P, true>(x);
  }
}

// Length-(TPR*P) FFT over a group of TPR lanes whose cross-lane step is one transpose through
// the group's LDS scratch instead of a DPP network (P a multiple of TPR, J = P / TPR):
//   X[k1 + P k2] = sum_t W_TPR^(t k2) W_L^(t k1) Y_t[k1],  Y_t = the in-register FFT of lane t.
// Lane t writes W_L^(t k1) Y_t[k1] to row t of the scratch, then owns J values of k1
// (k1 = J t + j) and transforms them over the TPR rows in registers.
//  forward: in  x[m] = data[TPR m + t];  out x[j TPR + k2] = DATA[(J t + j) + P k2]
//  inverse: in  x[j TPR + k2] = DATA[(J t + j) + P k2];  out x[m] = data[TPR m + t] (x TPR P)
// ``twl``: row t of the lane twiddle table, twl[k] = exp(-2 pi i t k / (TPR P)) for k < P
// (twl_row(P) elements per row, 16-byte aligned): the lane reads its twiddles at constant
// offsets from one base, two per 16-byte load.  ``xch``: TPR rows of xch_row(P) complex,
// private to the group.  The group lives in one wave, whose LDS accesses complete in issue
// order; the wavefront-scope fences keep the compiler from moving them across the exchange.
constexpr int xch_row(int P) { return P + 2; }  // row stride: 16-byte aligned pairs, rotated banks
constexpr int twl_row(int P) { return P + 2; }  // lane twiddle row stride (rotated banks)

// Row t of the lane twiddle table for length L = TPR P, entry k = exp(-2 pi i t k / L).
LX_DEV float2 twl_entry(int L, int t, int k) {
  double sn, cs;
  sincospi(2.0 * double((t * k) % L) / double(L), &sn, &cs);
  return make_float2((float)cs, (float)-sn);
}

template <int P>
LX_DEV void twl_mul(float2* x, const float2* twl) {
  static_assert(P % 2 == 0, "lane twiddle rows are read in pairs");
#pragma unroll
  for (int k = 0; k < P; k += 2) {
    const float4 w = *reinterpret_cast<const float4*>(twl + k);
    if (k > 0) x[k] = cmul(x[k], make_float2(w.x, w.y));
    x[k + 1] = cmul(x[k + 1], make_float2(w.z, w.w));
  }
}

template <int P>
LX_DEV void twl_mulc(float2* x, const float2* twl) {
#pragma unroll
  for (int k = 0; k < P; k += 2) {
    const float4 w = *reinterpret_cast<const float4*>(twl + k);
    if (k > 0) x[k] = cmulc(x[k], make_float2(w.x, w.y));
    x[k + 1] = cmulc(x[k + 1], make_float2(w.z, w.w));
  }
}

template <int P, int TPR, bool INV>
LX_DEV void fft_xlds(float2* x, int t, const float2* twl, float2* xch) {
  constexpr int J = P / TPR, RS = xch_row(P);
  static_assert(P % TPR == 0 && J >= 1, "fft_xlds needs P >= TPR");
  if constexpr (!INV) {
    fft_reg<P, false>(x);
    twl_mul<P>(x, twl);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
#pragma unroll
    for (int k = 0; k < P; k += 2) {
      const float4 v = make_float4(x[k].x, x[k].y, x[k + 1].x, x[k + 1].y);
      *reinterpret_cast<float4*>(&xch[t * RS + k]) = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
    for (int r = 0; r < TPR; ++r) {
      if constexpr (J % 2 == 0) {
#pragma unroll
        for (int j2 = 0; j2 < J; j2 += 2) {
          const float4 v = *reinterpret_cast<const float4*>(&xch[r * RS + J * t + j2]);
          x[j2 * TPR + r] = make_float2(v.x, v.y);
          x[(j2 + 1) * TPR + r] = make_float2(v.z, v.w);
        }
      } else {
#pragma unroll
        for (int j2 = 0; j2 < J; ++j2) x[j2 * TPR + r] = xch[r * RS + J * t + j2];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int j2 = 0; j2 < J; ++j2) fft_reg<TPR, false>(x + j2 * TPR);
  } else {
#pragma unroll
    for (int j2 = 0; j2 < J; ++j2) fft_reg<TPR, true>(x + j2 * TPR);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
#pragma unroll
    for (int r = 0; r < TPR; ++r) {
      if constexpr (J % 2 == 0) {
#pragma unroll
        for (int j2 = 0; j2 < J; j2 += 2) {
          const float4 v = make_float4(x[j2 * TPR + r].x, x[j2 * TPR + r].y, x[(j2 + 1) * TPR + r].x,
                                       x[(j2 + 1) * TPR + r].y);
          *reinterpret_cast<float4*>(&xch[r * RS + J * t + j2]) = v;
        }
      } else {
#pragma unroll
        for (int j2 = 0; j2 < J; ++j2) xch[r * RS + J * t + j2] = x[j2 * TPR + r];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
    for (int k = 0; k < P; k += 2) {
      const float4 v = *reinterpret_cast<const float4*>(&xch[t * RS + k]);
      x[k] = make_float2(v.x, v.y);
      x[k + 1] = make_float2(v.z, v.w);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    twl_mulc<P>(x, twl);
    fft_reg<P, true>(x);
  }
}

// ----------------------------------------------------------------- growth functions
// Ids follow the registry order of leniax/growth_functions.py:256-264.
enum : int {
  LX_GF_POLY_QUAD4 = 0,
  LX_GF_GAUSSIAN = 1,
  LX_GF_GAUSSIAN_TARGET = 2,
  LX_GF_STEP = 3,
  LX_GF_STAIRCASE = 4,
  LX_GF_TRIANGLE = 5,
  LX_GF_IDENTITY = 6,
};

LX_DEV float staircase01(float m, float s, float X) {
  float out = 0.f;
  out += (X >= m - s && X < m - s / 2.f) ? 0.5f : 0.f;
  out += (X >= m - s / 2.f && X <= m + s / 2.f) ? 1.f : 0.f;
  out += (X > m + s / 2.f && X <= m + s) ? 0.5f : 0.f;
  return out;
}

LX_DEV float triangle01(float m, float s, float X) {
  float left = m - s, right = m + s, out = 0.f;
  out += (X >= left && X < m) ? (X - left) / (m - left) : 0.f;
  out += (X >= m && X <= right) ? (X - right) / (m - right) : 0.f;
  return out;
}

// Products and sums that must round separately (no FMA contraction), as leniax's
// ``state + dt * field`` and ``w * g`` do.
LX_DEV float mul_nc(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
LX_DEV float add_nc(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}
LX_DEV float sub_nc(float a, float b) {
#pragma clang fp contract(off)
  return a - b;
}

// g(U) for a compile-time growth function id.
template <int KIND>
LX_DEV float growth_k(float m, float s, float U) {
  if constexpr (KIND == LX_GF_POLY_QUAD4) {  // leniax/growth_functions.py:6-45
    float d = U - m;
    float o = 1.f - (d * d) / (9.f * (s * s));
    o = fmaxf(0.f, o);
    o = o * o;
    o = o * o;
    return 2.f * o - 1.f;
  } else if constexpr (KIND == LX_GF_GAUSSIAN) {  // :48-82
    float z = (U - m) / s;
    return 2.f * __expf(-(z * z) / 2.f) - 1.f;
  } else if constexpr (KIND == LX_GF_GAUSSIAN_TARGET) {  // :85-119
    float z = (U - m) / s;
    return __expf(-(z * z) / 2.f);
  } else if constexpr (KIND == LX_GF_STEP) {  // :122-160
    return fabsf(U - m) <= s ? 1.f : -1.f;
  } else if constexpr (KIND == LX_GF_STAIRCASE) {  // :163-201
    return 2.f * staircase01(m, s, U) - 1.f;
  } else if constexpr (KIND == LX_GF_TRIANGLE) {  // :204-243
    return 2.f * triangle01(m, s, U) - 1.f;
  } else {  // identity :246-253
    return U;
  }
}

// g(U) with per-rule reciprocals precomputed: inv_s = 1/s, inv_9s2 = 1/(9 s^2)
// (one rounding away from leniax's divisions).
template <int KIND>
LX_DEV float growth_pre(float m, float s, float inv_s, float inv_9s2, float U) {
  if constexpr (KIND == LX_GF_POLY_QUAD4) {
    float d = U - m;
    float o = fmaxf(0.f, 1.f - (d * d) * inv_9s2);
    o = o * o;
    o = o * o;
    return 2.f * o - 1.f;
  } else if constexpr (KIND == LX_GF_GAUSSIAN) {
    float z = (U - m) * inv_s;
    return 2.f * __expf(-0.5f * (z * z)) - 1.f;
  } else if constexpr (KIND == LX_GF_GAUSSIAN_TARGET) {
    float z = (U - m) * inv_s;
    return __expf(-0.5f * (z * z));
  } else {
    return growth_k<KIND>(m, s, U);
  }
}

// g(U) by runtime id (plain path).
LX_DEV float growth(int kind, float m, float s, float U) {
  switch (kind) {
    case LX_GF_POLY_QUAD4: {  // leniax/growth_functions.py:6-45
      float d = U - m;
      float o = 1.f - (d * d) / (9.f * (s * s));
      o = fmaxf(0.f, o);
      o = o * o;
      o = o * o;
      return 2.f * o - 1.f;
    }
    case LX_GF_GAUSSIAN: {  // :48-82
      float z = (U - m) / s;
      return 2.f * __expf(-(z * z) / 2.f) - 1.f;
    }
    case LX_GF_GAUSSIAN_TARGET: {  // :85-119
      float z = (U - m) / s;
      return __expf(-(z * z) / 2.f);
    }
    case LX_GF_STEP:  // :122-160
      return fabsf(U - m) <= s ? 1.f : -1.f;
    case LX_GF_STAIRCASE:  // :163-201
      return 2.f * staircase01(m, s, U) - 1.f;
    case LX_GF_TRIANGLE:  // :204-243
      return 2.f * triangle01(m, s, U) - 1.f;
    default:  // identity :246-253
      return U;
  }
}

// g(U) by runtime id with the kernel's reciprocals precomputed (inv_s = 1/s, inv_9s2 = 1/(9 s^2)):
// the Gaussians and the quartic polynomial without a division per cell (one rounding away from
// leniax's divisions), the others as growth().
LX_DEV float growth_rcp(int kind, float m, float s, float inv_s, float inv_9s2, float U) {
  if (kind == LX_GF_GAUSSIAN || kind == LX_GF_GAUSSIAN_TARGET) {
    const float z = (U - m) * inv_s;
    const float e = __expf(-0.5f * (z * z));
    return kind == LX_GF_GAUSSIAN ? 2.f * e - 1.f : e;
  }
  if (kind == LX_GF_POLY_QUAD4) {
    const float d = U - m;
    float o = fmaxf(0.f, 1.f - (d * d) * inv_9s2);
    o = o * o;
    o = o * o;
    return 2.f * o - 1.f;
  }
  return growth(kind, m, s, U);
}

// dg/dU, dg/dm, dg/ds for the differentiable growth functions (adjoint).
// inv_s = 1/s and inv_9s2 = 1/(9 s^2) are the kernel's reciprocals (as growth_rcp takes them):
// the Gaussians and the quartic need no division per cell.
LX_DEV void growth_grad(int kind, float m, float s, float inv_s, float inv_9s2, float U, float& dU, float& dm,
                        float& ds) {
  switch (kind) {
    case LX_GF_POLY_QUAD4: {
      float d = U - m;
      float o = 1.f - (d * d) * inv_9s2;
      if (o <= 0.f) { dU = dm = ds = 0.f; return; }
      float o3 = o * o * o;
      float dgdo = 8.f * o3;
      dU = dgdo * (-2.f * d * inv_9s2);
      dm = -dU;
      ds = dgdo * (2.f * d * d * inv_9s2 * inv_s);
      return;
    }
    case LX_GF_GAUSSIAN:
    case LX_GF_GAUSSIAN_TARGET: {
      float z = (U - m) * inv_s;
      float e = __expf(-0.5f * (z * z));
      float amp = kind == LX_GF_GAUSSIAN ? 2.f : 1.f;
      dU = -amp * e * z * inv_s;
      dm = -dU;
      ds = amp * e * z * z * inv_s;
      return;
    }
    case LX_GF_TRIANGLE: {  // 2 (U - (m - s)) / s on [m - s, m), 2 (U - (m + s)) / (-s) on [m, m + s]
      const float d = U - m;
      if (U >= m - s && U < m) {
        dU = 2.f / s; dm = -dU; ds = -2.f * d / (s * s);
      } else if (U >= m && U <= m + s) {
        dU = -2.f / s; dm = -dU; ds = 2.f * d / (s * s);
      } else {
        dU = dm = ds = 0.f;
      }
      return;
    }
    case LX_GF_IDENTITY:
      dU = 1.f; dm = ds = 0.f; return;
    default:  // step, staircase: piecewise constant
      dU = dm = ds = 0.f; return;
  }
}

// ----------------------------------------------------------------- wave reductions
LX_DEV float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// wave_sum of N values at once, round by round: one value's exchange overlaps another's wait.
// Each value's additions are wave_sum's, in its order (the same sums bit for bit).
template <int N>
LX_DEV void wave_sum_n(float (&v)[N]) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] += __shfl_xor(v[j], off, 64);
  }
}

// Sum over the 64 lanes of a wave, complete in lane 63 only (the other lanes hold partial
// sums): DPP butterflies within each row of 16 lanes (quad permutes, half-row and row
// mirrors), then the row broadcasts of lanes 15 and 31.  No LDS traffic.
template <typename T>
LX_DEV T wave_sum_last(T v) {
  static_assert(sizeof(T) == 4, "32-bit values");
  auto mv = [](T a, int ctrl_sel) -> T {
    const int i = __builtin_bit_cast(int, a);
    int r;
    switch (ctrl_sel) {
      case 0: r = __builtin_amdgcn_update_dpp(0, i, LX_QP(1, 0, 3, 2), 0xF, 0xF, false); break;
      case 1: r = __builtin_amdgcn_update_dpp(0, i, LX_QP(2, 3, 0, 1), 0xF, 0xF, false); break;
      case 2: r = __builtin_amdgcn_update_dpp(0, i, 0x141, 0xF, 0xF, false); break;  // row_half_mirror
      case 3: r = __builtin_amdgcn_update_dpp(0, i, 0x140, 0xF, 0xF, false); break;  // row_mirror
      case 4: r = __builtin_amdgcn_update_dpp(0, i, 0x142, 0xA, 0xF, false); break;  // row_bcast:15 -> rows 1, 3
      default: r = __builtin_amdgcn_update_dpp(0, i, 0x143, 0xC, 0xF, false); break; // row_bcast:31 -> rows 2, 3
    }
    return __builtin_bit_cast(T, r);
  };
  v += mv(v, 0);
  v += mv(v, 1);
  v += mv(v, 2);
  v += mv(v, 3);
  v += mv(v, 4);
  v += mv(v, 5);
  return v;
}

// wave_sum_last of N values at once, stage by stage: a stage of one value fills the wait states
// that follow the previous stage of another.  Each value's additions are wave_sum_last's, in its
// order, so the sums are the same bit for bit.
template <int N, typename T>
LX_DEV void wave_sum_last_n(T (&v)[N]) {
  static_assert(sizeof(T) == 4, "32-bit values");
  auto mv = [](T a, int ctrl_sel) -> T {
    const int i = __builtin_bit_cast(int, a);
    int r;
    switch (ctrl_sel) {
      case 0: r = __builtin_amdgcn_update_dpp(0, i, LX_QP(1, 0, 3, 2), 0xF, 0xF, false); break;
      case 1: r = __builtin_amdgcn_update_dpp(0, i, LX_QP(2, 3, 0, 1), 0xF, 0xF, false); break;
      case 2: r = __builtin_amdgcn_update_dpp(0, i, 0x141, 0xF, 0xF, false); break;
      case 3: r = __builtin_amdgcn_update_dpp(0, i, 0x140, 0xF, 0xF, false); break;
      case 4: r = __builtin_amdgcn_update_dpp(0, i, 0x142, 0xA, 0xF, false); break;
      default: r = __builtin_amdgcn_update_dpp(0, i, 0x143, 0xC, 0xF, false); break;
    }
    return __builtin_bit_cast(T, r);
  };
#pragma unroll
  for (int stage = 0; stage < 6; ++stage) {
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] += mv(v[j], stage);
  }
}
