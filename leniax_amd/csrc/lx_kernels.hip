// Kernels and their spectra built on the device, in float64 (leniax/kernels.py:66-158 FFT
// branch, the shapes circle_2d / ellipse_2d / oriented_ellipse_2d of :176-309 and the kernel
// functions of leniax/kernel_functions.py:7-250), for many kernels at once: the per-individual
// host build of a quality-diversity loop becomes one launch.
//
// A kernel of radius kr = ceil(r R) lives on a 2kr x 2kr grid, centre kr, coordinates
// c = (index - kr) / (r R); it is zero-padded to the world with (world - 2kr) // 2 cells
// before (leniax/kernels.py:91-101) and fftshift-ed (numpy.fft.fftshift: cell i moves to
// (i + n // 2) mod n), so that its real FFT is leniax's K = fftn(fftshift(k)).
#include <algorithm>

#include "lx_internal.h"

namespace {

constexpr int kThreads = 256;

// kernel functions on [0, 1] in float64 (leniax/kernel_functions.py), ids of enum lx_kernel_fn
__device__ double kf_eval(int kf, const double* q, double X) {
  switch (kf) {
    case LX_KF_POLY_QUAD: {  // (4 X (1 - X))^q  :7-37
      const double o = 4.0 * X * (1.0 - X);
      return pow(o, q[0]);
    }
    case LX_KF_GAUSS_BUMP:  // exp(q (q - 1 / (X (1 - X) + eps)))  :40-69
      return exp(q[0] * (q[0] - 1.0 / (X * (1.0 - X) + 1e-7)));
    case LX_KF_STEP:  // 1 on [q, 1 - q]  :72-102
      return (X >= q[0] && X <= 1.0 - q[0]) ? 1.0 : 0.0;
    case LX_KF_GAUSS: {  // exp(-((X - q) / (0.3 q))^2 / 2)  :105-134
      const double z = (X - q[0]) / (0.3 * q[0]);
      return exp(-(z * z) / 2.0);
    }
    case LX_KF_THRESHOLD:  // 1 for X >= q  :137-167
      return X >= q[0] ? 1.0 : 0.0;
    case LX_KF_STAIRCASE: {  // :170-208
      const double m = q[0], s = q[1];
      double out = (X >= m - s && X < m - s / 2.0) ? 0.5 : 0.0;
      out += (X >= m - s / 2.0 && X <= m + s / 2.0) ? 1.0 : 0.0;
      out += (X > m + s / 2.0 && X <= m + s) ? 0.5 : 0.0;
      return out;
    }
    default: {  // triangle :211-250
      const double m = q[0], s = q[1], left = m - s, right = m + s;
      double out = (X >= left && X < m) ? (X - left) / (m - left) : 0.0;
      out += (X >= m && X <= right) ? (X - right) / (m - right) : 0.0;
      return out;
    }
  }
}

// Unnormalised value of cell (i, j) of the 2kr x 2kr grid, and the factor its normalisation
// divides by the sum of (circle, ellipse: the value; oriented ellipse: |value|).  ring >= 0:
// only that ring, with weight 1 (the ring basis of the b gradient).
struct CellVal {
  double v, norm_term, post;  // kernel = v / sum(norm_term) * post
};

__device__ CellVal cell_value(const lx_kernel_desc& d, double R, int i, int j, int ring) {
  const double rR = d.r * R;
  const double c0 = double(i - d.kr) / rR, c1 = double(j - d.kr) / rR;
  double x0 = c0, x1 = c1, dist;
  if (d.shape == LX_KSHAPE_CIRCLE) {
    dist = sqrt(c0 * c0 + c1 * c1);
  } else {  // leniax/kernels.py:236-243 rotated coordinates, axes a and b
    const double th = d.theta * 3.14159265358979323846;
    const double ct = cos(th), st = sin(th);
    x0 = c0 * ct + c1 * st;
    x1 = -c0 * st + c1 * ct;
    const double u = x0 / d.a, w = x1 / d.b;
    dist = sqrt(u * u + w * w);
  }
  const double Bd = double(d.nb) * dist;
  int bi = int(floor(Bd));
  bi = bi < d.nb - 1 ? bi : d.nb - 1;
  const double X = Bd - floor(Bd);  // B_dist % 1 (B_dist >= 0)
  double b = d.bs[bi];
  if (ring >= 0) b = (bi == ring) ? 1.0 : 0.0;
  const double v = (dist < 1.0) ? kf_eval(d.kf, d.kf_params, X) * b : 0.0;
  CellVal out;
  if (d.shape == LX_KSHAPE_ORIENTED_ELLIPSE) {  // :303-305 times x0, normalised by sum |.|
    out.v = v * x0;
    out.norm_term = fabs(out.v);
    out.post = 1.0;
  } else {
    out.v = v;
    out.norm_term = v;
    out.post = 1.0;
    if (d.shape == LX_KSHAPE_ELLIPSE)  // :256-258 sign of x0 outside (-0.01, 0.01)
      out.post = x0 < -0.01 ? -1.0 : (x0 > 0.01 ? 1.0 : x0);
  }
  return out;
}

__device__ double block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < kThreads / 64; ++k) t += red[k];  // fixed order: deterministic
  __syncthreads();
  return t;
}

struct BuildArgs {
  const lx_kernel_desc* desc;
  int64_t n;       // descriptors
  int H, W;        // world
  double R;
  float* out;      // [n * per][H][W]
  int basis;       // 0: the kernels; 1: the unnormalised ring basis, LX_MAX_RINGS arrays per desc
};

__global__ void zero_kernel(float* p, int64_t count) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = 0.f;
}

// grid: (descriptors, rings when basis).  Pass 1 sums the normalisation terms in float64, pass 2
// writes the rounded values at their padded, shifted places.
__global__ __launch_bounds__(kThreads) void build_kernel_kernel(BuildArgs a) {
  __shared__ double red[kThreads / 64];
  const int64_t k = blockIdx.x;
  const int ring = a.basis ? int(blockIdx.y) : -1;
  const lx_kernel_desc d = a.desc[k];
  const int ks = 2 * d.kr;
  float* out = a.out + (a.basis ? (k * LX_MAX_RINGS + ring) : k) * int64_t(a.H) * a.W;
  // a descriptor that does not fit leaves its slot zero (the host checks them before the launch)
  if (d.kr < 1 || ks > a.H || ks > a.W || d.nb < 1 || d.nb > LX_MAX_RINGS) return;
  if (a.basis && ring >= d.nb) return;  // unused ring slots stay zero
  const int ph = (a.H - ks) / 2, pw = (a.W - ks) / 2;
  double s = 0.0;
  for (int c = threadIdx.x; c < ks * ks; c += kThreads) s += cell_value(d, a.R, c / ks, c % ks, ring).norm_term;
  const double total = a.basis ? 1.0 : block_sum(s, red);
  for (int c = threadIdx.x; c < ks * ks; c += kThreads) {
    const int i = c / ks, j = c % ks;
    const CellVal cv = cell_value(d, a.R, i, j, ring);
    const double v = cv.v / total * cv.post;
    const int y = (i + ph + a.H / 2) % a.H, x = (j + pw + a.W / 2) % a.W;
    out[int64_t(y) * a.W + x] = float(v);
  }
}

}  // namespace

extern "C" int lx_build_kernels(int32_t ndim, const int32_t* dims, const lx_kernel_desc* desc_d, int64_t n,
                                double R, int32_t max_kr, int32_t basis, float* out_d, void* stream) {
  if (!dims || n < 0) return lx_fail(LX_EINVAL, "bad arguments");
  if (n == 0) return LX_OK;
  if (!desc_d || !out_d) return lx_fail(LX_EINVAL, "NULL pointer");
  if (ndim != 2) return lx_fail(LX_ENOTSUP, "lx_build_kernels: 2-D kernels (3-D worlds pass theirs as raw)");
  const int H = dims[0], W = dims[1];
  if (H < 1 || W < 1) return lx_fail(LX_EINVAL, "world %dx%d", H, W);
  if (max_kr < 1 || 2 * max_kr > H || 2 * max_kr > W)
    return lx_fail(LX_EINVAL, "kernel radius %d does not fit the %dx%d world", max_kr, H, W);
  if (n > 2147483647) return lx_fail(LX_ENOTSUP, "too many kernels");
  hipStream_t st = (hipStream_t)stream;
  const int64_t count = n * (basis ? LX_MAX_RINGS : 1) * int64_t(H) * W;
  hipLaunchKernelGGL(zero_kernel, dim3(unsigned(std::min<int64_t>((count + 255) / 256, 65536))), dim3(256), 0, st,
                     out_d, count);
  BuildArgs a;
  a.desc = desc_d; a.n = n; a.H = H; a.W = W; a.R = R; a.out = out_d; a.basis = basis ? 1 : 0;
  hipLaunchKernelGGL(build_kernel_kernel, dim3(unsigned(n), basis ? LX_MAX_RINGS : 1), dim3(kThreads), 0, st, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LX_OK : lx_fail(LX_EDEVICE, "kernel build launch: %s", hipGetErrorString(e));
}

extern "C" int lx_build_spectra(int32_t ndim, const int32_t* dims, const lx_kernel_desc* desc_d, int64_t n, double R,
                                int32_t max_kr, void* K_half_d, float* work_d, void* stream) {
  if (n == 0) return LX_OK;
  if (!K_half_d || !work_d) return lx_fail(LX_EINVAL, "NULL pointer");
  if (int rc = lx_build_kernels(ndim, dims, desc_d, n, R, max_kr, 0, work_d, stream)) return rc;
  return lx_rfft2(ndim, dims, n, work_d, K_half_d, stream);
}
