// Initial states on the device (leniax/initializations.py:10-116, leniax/perlin.py:17-71, the
// quantisation of leniax/loader.py:16-30), from jax.random's counter-based stream.
//
// jax.random draws from the Threefry-2x32 block cipher (20 rounds; Salmon et al., "Parallel
// random numbers: as easy as 1, 2, 3", SC 2011).  The words of a request of n values are the
// cipher of the counts 0..n-1 split in two halves (a zero appended when n is odd): word i < h is
// the first output of the pair (i, h + i), word h + i its second output, h = ceil(n / 2).
// Uniform floats take the top 23 bits as the mantissa of a number in [1, 2), minus 1.
#include <algorithm>

#include "lx_internal.h"

namespace {

LX_DEV uint32_t rotl32(uint32_t x, int d) { return (x << d) | (x >> (32 - d)); }

// Threefry-2x32, 20 rounds: rotations 13 15 26 6 / 17 29 16 24, a key injection every 4 rounds
// (key schedule k0, k1, k0 ^ k1 ^ 0x1BD11BDA, plus the injection count on the second word).
LX_DEV void threefry2x32(uint32_t k0, uint32_t k1, uint32_t& x0, uint32_t& x1) {
  const uint32_t ks[3] = {k0, k1, k0 ^ k1 ^ 0x1BD11BDAu};
  constexpr int rot[2][4] = {{13, 15, 26, 6}, {17, 29, 16, 24}};
  x0 += ks[0];
  x1 += ks[1];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      x0 += x1;
      x1 = rotl32(x1, rot[i & 1][r]);
      x1 ^= x0;
    }
    x0 += ks[(i + 1) % 3];
    x1 += ks[(i + 2) % 3] + uint32_t(i + 1);
  }
}

// word i of jax.random's stream of n words under key (k0, k1)
struct Stream {
  uint32_t k0, k1;
  int64_t n, h;
};

LX_DEV float unit_float(uint32_t bits) {  // [0, 1): mantissa bits of [1, 2) minus 1
  return __uint_as_float((bits >> 9) | 0x3f800000u) - 1.f;
}

__global__ void bits_kernel(Stream s, uint32_t* out) {
  for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < s.h; j += int64_t(gridDim.x) * blockDim.x) {
    uint32_t x0 = uint32_t(j), x1 = (s.h + j < s.n) ? uint32_t(s.h + j) : 0u;
    threefry2x32(s.k0, s.k1, x0, x1);
    out[j] = x0;
    if (s.h + j < s.n) out[s.h + j] = x1;
  }
}

// leniax/loader.py:16-30: round to the NB_CHARS^2 - 1 grid (half to even, as numpy's round)
LX_DEV float quantise(float v, float max_val) { return float(int(rintf(v * max_val))) / max_val; }

// random_uniform (leniax/initializations.py:10-33): value c of noise b (cells per noise: cells)
// is max(0, u * maxval_b + 0) for the stream word b * cells + c, then quantised.
__global__ void uniform_kernel(Stream s, int64_t cells, const float* maxval, float max_q, float* out) {
  for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < s.h; j += int64_t(gridDim.x) * blockDim.x) {
    uint32_t x0 = uint32_t(j), x1 = (s.h + j < s.n) ? uint32_t(s.h + j) : 0u;
    threefry2x32(s.k0, s.k1, x0, x1);
    const float m0 = maxval[j / cells];
    out[j] = quantise(fmaxf(0.f, unit_float(x0) * m0 + 0.f), max_q);
    if (s.h + j < s.n) {
      const float m1 = maxval[(s.h + j) / cells];
      out[s.h + j] = quantise(fmaxf(0.f, unit_float(x1) * m1 + 0.f), max_q);
    }
  }
}

// angles of the perlin gradients: 2 pi u (leniax/initializations.py:61-63), u from the stream
__global__ void angles_kernel(Stream s, float* out) {
  for (int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; j < s.h; j += int64_t(gridDim.x) * blockDim.x) {
    uint32_t x0 = uint32_t(j), x1 = (s.h + j < s.n) ? uint32_t(s.h + j) : 0u;
    threefry2x32(s.k0, s.k1, x0, x1);
    constexpr float two_pi = 6.28318530717958647692f;
    out[j] = two_pi * unit_float(x0);
    if (s.h + j < s.n) out[s.h + j] = two_pi * unit_float(x1);
  }
}

struct PerlinArgs {
  const float* angles;  // [nb][r0][r1]
  const float* scaling; // [nb]
  float* out;           // [nb][H][W]
  int H, W, r0, r1;
  float max_q;
  int quantise;
};

// Gradient noise of leniax/perlin.py:17-71 at cell (y, x) of noise b: the lattice has r0 x r1
// cells of d0 = H / r0 by d1 = W / r1 cells (integer division), gradient (cos a, sin a) at each
// lattice point, wrapped; the corner gradients of a cell are those the reference's repeated and
// offset gradient arrays place there (rows y and y + (r0 + 1) d0 - H of the repeated array, each
// standing for lattice row (row / d0) mod r0; columns alike); local coordinates
// t = (y r0 / H) mod 1, fade f(t) = t^3 (t (6 t - 15) + 10), result sqrt(2) times the blend.
LX_DEV float perlin_at(const PerlinArgs& a, const float* ang, int y, int x) {
  const int d0 = a.H / a.r0, d1 = a.W / a.r1;
  const int o0 = (a.r0 + 1) * d0 - a.H, o1 = (a.r1 + 1) * d1 - a.W;
  const int i0 = (y / d0) % a.r0, i1 = ((y + o0) / d0) % a.r0;
  const int j0 = (x / d1) % a.r1, j1 = ((x + o1) / d1) % a.r1;
  // the ramp y r0 / H in float64 (numpy's arange of the static grid), then float32 mod 1
  float t0 = float(double(y) * (double(a.r0) / double(a.H)));
  float t1 = float(double(x) * (double(a.r1) / double(a.W)));
  t0 -= floorf(t0);
  t1 -= floorf(t1);
  float sn, cs;
  sincosf(ang[i0 * a.r1 + j0], &sn, &cs);
  const float n00 = t0 * cs + t1 * sn;
  sincosf(ang[i1 * a.r1 + j0], &sn, &cs);
  const float n10 = (t0 - 1.f) * cs + t1 * sn;
  sincosf(ang[i0 * a.r1 + j1], &sn, &cs);
  const float n01 = t0 * cs + (t1 - 1.f) * sn;
  sincosf(ang[i1 * a.r1 + j1], &sn, &cs);
  const float n11 = (t0 - 1.f) * cs + (t1 - 1.f) * sn;
  const float f0 = t0 * t0 * t0 * (t0 * (t0 * 6.f - 15.f) + 10.f);
  const float f1 = t1 * t1 * t1 * (t1 * (t1 * 6.f - 15.f) + 10.f);
  const float n0 = n00 * (1.f - f0) + f0 * n10;
  const float n1 = n01 * (1.f - f0) + f0 * n11;
  return 1.41421356237309504880f * ((1.f - f1) * n0 + f1 * n1);
}

__device__ float block_reduce(float v, bool is_max, float* red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float o = __shfl_xor(v, off, 64);
    v = is_max ? fmaxf(v, o) : fminf(v, o);
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  float t = red[0];
  for (int k = 1; k < int(blockDim.x) / 64; ++k) t = is_max ? fmaxf(t, red[k]) : fminf(t, red[k]);
  __syncthreads();
  return t;
}

// leniax/initializations.py:36-80: noise b shifted to minimum 0, divided by its maximum, times
// scaling_b, quantised.  One block per noise; the noise is recomputed in each pass.
__global__ __launch_bounds__(256) void perlin_kernel(PerlinArgs a) {
  __shared__ float red[4];
  const int b = blockIdx.x;
  const float* ang = a.angles + int64_t(b) * a.r0 * a.r1;
  const int cells = a.H * a.W;
  float mn = INFINITY;
  for (int c = threadIdx.x; c < cells; c += 256) mn = fminf(mn, perlin_at(a, ang, c / a.W, c % a.W));
  mn = block_reduce(mn, false, red);
  float mx = -INFINITY;
  for (int c = threadIdx.x; c < cells; c += 256) mx = fmaxf(mx, perlin_at(a, ang, c / a.W, c % a.W) - mn);
  mx = block_reduce(mx, true, red);
  const float sc = a.scaling[b];
  float* out = a.out + int64_t(b) * cells;
  for (int c = threadIdx.x; c < cells; c += 256) {
    const float v = (perlin_at(a, ang, c / a.W, c % a.W) - mn) / mx * sc;
    out[c] = a.quantise ? quantise(v, a.max_q) : v;
  }
}

Stream make_stream(const uint32_t* key, int64_t n) {
  Stream s;
  s.k0 = key[0];
  s.k1 = key[1];
  s.n = n;
  s.h = (n + 1) / 2;
  return s;
}

unsigned grid_for(int64_t h) { return unsigned(std::max<int64_t>(1, std::min<int64_t>((h + 255) / 256, 8192))); }

}  // namespace

extern "C" int lx_random_bits(const uint32_t* key, int64_t n, uint32_t* out_d, void* stream) {
  if (n < 0) return lx_fail(LX_EINVAL, "bad arguments");
  if (n == 0) return LX_OK;
  if (!key || !out_d) return lx_fail(LX_EINVAL, "NULL pointer");
  if (n > 0xFFFFFFFFll) return lx_fail(LX_ENOTSUP, "at most 2^32 words per request");
  const Stream s = make_stream(key, n);
  hipLaunchKernelGGL(bits_kernel, dim3(grid_for(s.h)), dim3(256), 0, (hipStream_t)stream, s, out_d);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LX_OK : lx_fail(LX_EDEVICE, "random bits launch: %s", hipGetErrorString(e));
}

extern "C" int lx_init_uniform(const uint32_t* key, int64_t nb, int64_t cells, const float* maxval_d, float* out_d,
                               void* stream) {
  if (nb < 0 || cells < 1) return lx_fail(LX_EINVAL, "bad arguments");
  if (nb == 0) return LX_OK;
  if (!key || !maxval_d || !out_d) return lx_fail(LX_EINVAL, "NULL pointer");
  if (nb * cells > 0xFFFFFFFFll) return lx_fail(LX_ENOTSUP, "at most 2^32 cells per request");
  const Stream s = make_stream(key, nb * cells);
  hipLaunchKernelGGL(uniform_kernel, dim3(grid_for(s.h)), dim3(256), 0, (hipStream_t)stream, s, cells, maxval_d,
                     float(LX_NB_CHARS * LX_NB_CHARS - 1), out_d);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LX_OK : lx_fail(LX_EDEVICE, "uniform init launch: %s", hipGetErrorString(e));
}

extern "C" int lx_init_perlin(const uint32_t* key, int32_t nb, int32_t H, int32_t W, int32_t r0, int32_t r1,
                              const float* scaling_d, int32_t quantise, float* angles_d, float* out_d, void* stream) {
  if (nb < 0) return lx_fail(LX_EINVAL, "bad arguments");
  if (H < 1 || W < 1 || r0 < 1 || r1 < 1 || r0 > H || r1 > W)
    return lx_fail(LX_EINVAL, "perlin lattice %dx%d over a %dx%d world", r0, r1, H, W);
  if (nb == 0) return LX_OK;
  if (!key || !scaling_d || !angles_d || !out_d) return lx_fail(LX_EINVAL, "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  const Stream s = make_stream(key, int64_t(nb) * r0 * r1);
  hipLaunchKernelGGL(angles_kernel, dim3(grid_for(s.h)), dim3(256), 0, st, s, angles_d);
  PerlinArgs a;
  a.angles = angles_d; a.scaling = scaling_d; a.out = out_d;
  a.H = H; a.W = W; a.r0 = r0; a.r1 = r1;
  a.max_q = float(LX_NB_CHARS * LX_NB_CHARS - 1);
  a.quantise = quantise ? 1 : 0;
  hipLaunchKernelGGL(perlin_kernel, dim3(unsigned(nb)), dim3(256), 0, st, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LX_OK : lx_fail(LX_EDEVICE, "perlin init launch: %s", hipGetErrorString(e));
}
