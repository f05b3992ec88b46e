/* leniax_amd -- C ABI of the MI355X (gfx950) Lenia FFT stepping engine.
 *
 * Every pointer argument named *_d is a device pointer (HBM); the caller owns
 * all buffers.  Calls are asynchronous on the HIP stream passed in ("stream"
 * may be NULL for the default stream).  Return value: LX_OK (0) or a negative
 * errno-style code; lx_last_error() gives a thread-local message.  Array
 * layouts are leniax's channel-first ones: state [B][C][dims...] float32.
 * An empty batch (B == 0, n == 0) is valid: the call checks its sizes and
 * descriptors, launches nothing and returns LX_OK; its device pointers may
 * then be NULL.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the leniax repository).
 */
#ifndef LENIAX_AMD_H
#define LENIAX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LX_OK 0
#define LX_EINVAL (-22)
#define LX_ENOTSUP (-95)
#define LX_EDEVICE (-5)
#define LX_ENOMEM (-12)

#define LX_MAX_K 16 /* kernels (potential channels) per rule */
#define LX_MAX_C 4  /* channels per world */
#define LX_NSTAT_BASE 11 /* statistics per step before the C channel masses */

/* growth functions, registry order of leniax/growth_functions.py:256-264 */
enum lx_growth { LX_POLY_QUAD4 = 0, LX_GAUSSIAN = 1, LX_GAUSSIAN_TARGET = 2, LX_STEP = 3,
                 LX_STAIRCASE = 4, LX_TRIANGLE = 5, LX_IDENTITY = 6 };
/* state functions, leniax/core.py:322-326 */
enum lx_state_fn { LX_STATE_V1 = 0, LX_STATE_V2 = 1, LX_STATE_SIMPLE = 2 };
/* stop policies of lx_run */
enum lx_stop { LX_STOP_NONE = 0,   /* run every step (leniax run_scan*), N from the heuristics   */
               LX_STOP_LENIAX = 1, /* retire a world once leniax's heuristics fail (and t >= min_steps) */
               LX_STOP_MASS = 2 }; /* retire when the mass is 0 or covers > 90% of the grid */

/* kernel shapes, leniax/kernels.py:312-317 register (raw kernels are passed as arrays) */
enum lx_kernel_shape { LX_KSHAPE_CIRCLE = 0, LX_KSHAPE_ELLIPSE = 1, LX_KSHAPE_ORIENTED_ELLIPSE = 2 };
/* kernel functions, registry order of leniax/kernel_functions.py:253-261 */
enum lx_kernel_fn { LX_KF_POLY_QUAD = 0, LX_KF_GAUSS_BUMP = 1, LX_KF_GAUSS = 2, LX_KF_STEP = 3,
                    LX_KF_THRESHOLD = 4, LX_KF_STAIRCASE = 5, LX_KF_TRIANGLE = 6 };
#define LX_MAX_RINGS 8 /* ring weights b per kernel */

/* One kernel of leniax's kernels_params (leniax/kernels.py:176-309): k_params = [r, b, a, b_axis,
 * theta] and the kernel function's kf_params.  128 bytes, float64 parameters. */
typedef struct lx_kernel_desc {
  int32_t shape;             /* enum lx_kernel_shape */
  int32_t kf;                /* enum lx_kernel_fn */
  int32_t nb;                /* rings, 1..LX_MAX_RINGS */
  int32_t kr;                /* math.ceil(r * R): the kernel grid is 2 kr x 2 kr */
  double r;                  /* k_params[0] */
  double a, b;               /* ellipse axes k_params[2], k_params[3] (unused by circles) */
  double theta;              /* k_params[4]; the rotation angle is theta * pi */
  double kf_params[2];       /* kf_params (q, or m and s) */
  double bs[LX_MAX_RINGS];   /* ring weights k_params[1] */
} lx_kernel_desc;

/* The static part of leniax's update closure (leniax/helpers.py:401-515):
 * world shape, channel mapping, growth function per kernel, weighting, state fn. */
typedef struct lx_rule {
  int32_t ndim;             /* 2 or 3 world dimensions */
  int32_t dims[3];          /* world size; for ndim == 2 only dims[0..1] are used */
  int32_t C;                /* channels */
  int32_t K;                /* kernels == potential channels after leniax's tc_indices filter */
  int32_t cin[LX_MAX_K];    /* input channel of kernel k (kernels sorted by c_in, leniax/kernels.py:90) */
  int32_t gf[LX_MAX_K];     /* enum lx_growth of kernel k */
  int32_t state_fn;         /* enum lx_state_fn */
  int32_t average;          /* 1: weighted_mean (leniax/core.py:228-242), 0: weighted_sum (:202-225) */
  /* 0: K_half_d arguments hold Hermitian half spectra (lx_prepare_spectra layout, complex64).
   * 1 (2-D rules only): they hold real quadrants (lx_quadrant_spectra layout, float32
   * [S][K][dims[0]/2 + 1][dims[1]/2 + 1], k_bstride in float32 elements), the spectra of kernels
   * that are even along each axis (leniax's radially symmetric kernels).  Accepted by lx_update,
   * lx_update_with_spectra, lx_step_stats, lx_potential and lx_run (its multi-pass path); the
   * adjoint entry points and lx_run_static return LX_ENOTSUP. */
  int32_t k_quadrant;
} lx_rule;

/* Library / device information. */
const char* lx_last_error(void);
int lx_version(void);

/* Hermitian half spectra from leniax's full spectra.
 * Replaces the spectrum convention of leniax/kernels.py:145-149 (K = fftn(fftshift(k)))
 * as consumed by leniax/core.py:80-87.
 * K_full_d: complex64 [S][C][maxk][dims...]; slot[k] = c*maxk + j is the full-spectrum slot of
 * kernel k (leniax's tc_indices, leniax/core.py:97-98).
 * K_half_d: complex64 [S][K][dims[0..nd-2]][dims[nd-1]/2 + 1], holding (K(f) + conj(K(-f)))/2,
 * so that the real-to-complex path reproduces real(ifftn(fftn(A) * K)) for any K. */
int lx_prepare_spectra(const lx_rule* rule, const void* K_full_d, int64_t S, int32_t maxk,
                       const int32_t* slot, void* K_half_d, void* stream);

/* Real quadrants of half spectra for rules with k_quadrant = 1: K_quad[i][u][v] = Re K_half[i][u][v]
 * for u <= dims[0]/2, i < n (no counterpart in leniax, which keeps full complex spectra: leniax's
 * kernels are radially symmetric, so their spectra are real and even, and the quadrant holds a
 * quarter of the half spectrum's bytes).  rule: 2-D, dims[0] even.  K_half_d complex64
 * [n][dims[0]][dims[1]/2 + 1]; K_quad_d float32 [n][dims[0]/2 + 1][dims[1]/2 + 1]; dev_d float32[2]
 * receives the largest departure from a real even spectrum (max over i, u, v of |Im K[u][v]| +
 * |Im K[dims[0] - u][v]| + |Re K[u][v] - Re K[dims[0] - u][v]|, infinite for NaN) and the largest
 * |Re K|: the caller decides from them whether the kernels are even enough for the quadrant layout.
 * n == 0 writes nothing. */
int lx_quadrant_spectra(const lx_rule* rule, const void* K_half_d, int64_t n, void* K_quad_d, float* dev_d,
                        void* stream);

/* leniax/kernels.py:145-149: K = fftn(fftshift(k)) of real kernels on the device.
 * k_d float32 [n][dims] holds kernels already zero-padded to the world and fftshift-ed
 * (centre at index 0); K_full_d complex64 [n][dims]; work_d: n * prod(dims[:-1]) * (dims[-1]/2+1)
 * complex64 elements. */
int lx_kernel_spectrum(int32_t ndim, const int32_t* dims, int64_t n, const float* k_d, void* K_full_d,
                       void* work_d, void* stream);

/* leniax/kernels.py:66-149 for many kernels at once, on the device in float64: the kernels of
 * desc_d [n] (device memory) at world radius R, each normalised (sum 1; oriented ellipses: sum of
 * |k| 1), rounded to float32, zero-padded to the H x W world ((H - 2 kr) // 2 cells before) and
 * fftshift-ed, into out_d [n][H][W].  max_kr bounds every desc's kr (2 max_kr <= H, W).  With
 * basis != 0: per desc LX_MAX_RINGS unnormalised rings instead (ring i with weight 1, the others
 * 0; rings >= nb stay zero), out_d [n][LX_MAX_RINGS][H][W], the basis of the ring-weight gradient. */
int lx_build_kernels(int32_t ndim, const int32_t* dims, const lx_kernel_desc* desc_d, int64_t n, double R,
                     int32_t max_kr, int32_t basis, float* out_d, void* stream);
/* lx_build_kernels then their real FFTs: K_half_d [n][H][W/2+1] complex64, the half spectra of
 * leniax's K = fftn(fftshift(k)) (the lx_prepare_spectra layout for real kernels);
 * work_d: n * H * W floats. */
int lx_build_spectra(int32_t ndim, const int32_t* dims, const lx_kernel_desc* desc_d, int64_t n, double R,
                     int32_t max_kr, void* K_half_d, float* work_d, void* stream);

/* jax.random's stream (leniax's rng keys, leniax/utils.py:612-632): the Threefry-2x32 cipher
 * (20 rounds) under key[2] (host memory) of the counts 0..n-1 split in halves, as
 * jax.random.bits / uniform lay them out: out_d [n] uint32. */
int lx_random_bits(const uint32_t* key, int64_t n, uint32_t* out_d, void* stream);
/* leniax/initializations.py:10-33 random_uniform: nb states of `cells` cells, value c of state b
 * max(0, u maxval_d[b]) with u the stream word b * cells + c as a float in [0, 1), quantised
 * (leniax/loader.py:16-30).  key[2] is the subkey leniax draws from (host memory). */
int lx_init_uniform(const uint32_t* key, int64_t nb, int64_t cells, const float* maxval_d, float* out_d,
                    void* stream);
/* leniax/initializations.py:36-80 + leniax/perlin.py:17-71: nb perlin noises of H x W over an
 * r0 x r1 lattice, gradient angles 2 pi u from the stream of key[2] (host memory; the angles land in
 * angles_d [nb][r0][r1]), each shifted to minimum 0, divided by its maximum, times scaling_d[b],
 * quantised when quantise != 0: out_d [nb][H][W]. */
int lx_init_perlin(const uint32_t* key, int32_t nb, int32_t H, int32_t W, int32_t r0, int32_t r1,
                   const float* scaling_d, int32_t quantise, float* angles_d, float* out_d, void* stream);

/* Workspace bytes lx_potential / lx_update / lx_step_stats need for a batch of B worlds. */
int lx_workspace_size(const lx_rule* rule, int64_t B, size_t* bytes);

/* leniax/core.py:52-102 get_potential_fft: potential_d [B][K][dims] = real(ifftn(fftn(A) * K)).
 * K_half_d is indexed by world: spectrum of world b starts at b * k_bstride complex elements
 * (k_bstride == 0: one shared spectrum). */
int lx_potential(const lx_rule* rule, const float* state_d, int64_t B, const void* K_half_d,
                 int64_t k_bstride, float* potential_d, void* work_d, void* stream);

/* leniax/core.py:105-146 get_potential (the direct-convolution potential, helpers' fft=False):
 * potential_d[b][k][y][x] = sum_{i<kh, j<kw} state[b][cin k][(y+i-kh/2) mod H][(x+j-kw/2) mod W] * kern_d[k][i][j],
 * the depthwise cross-correlation of jax.lax.conv_general_dilated on the wrap-padded state
 * (padding of leniax/helpers.py:465-481).  kern_d: the K true kernels [K][kh][kw] (leniax's
 * [C*max_k, 1, kh, kw] without its padding slots).  2-D worlds of any size; kh, kw <= 64. */
int lx_potential_direct(const lx_rule* rule, const float* state_d, int64_t B, const float* kern_d, int32_t kh,
                        int32_t kw, float* potential_d, void* stream);

/* Reverse of lx_potential_direct (what jax.grad gives through leniax/core.py:105-146): with
 * dpot_d = dL/dpotential [B][K][H][W], dstate_d [B][C][H][W] = dL/dstate (overwritten: for each
 * channel the correlation of its kernels' dpot with the flipped kernels) and dkern_d [K][kh][kw] =
 * dL/dkernel summed over worlds (overwritten).  Either output may be NULL. */
int lx_potential_direct_adjoint(const lx_rule* rule, const float* state_d, int64_t B, const float* kern_d,
                                int32_t kh, int32_t kw, const float* dpot_d, float* dstate_d, float* dkern_d,
                                void* stream);

/* leniax/core.py:163-199 get_field (+ weighted_sum/weighted_mean):
 * field_d [B][C][dims] = sum_k w[c][k] g_k(potential_k) (/ sum_k w[c][k]).
 * gf_d [K][2] (m, s) and w_d [C][K] per world with strides gf_bstride / w_bstride floats (0: shared). */
int lx_field(const lx_rule* rule, const float* potential_d, int64_t B, const float* gf_d, int64_t gf_bstride,
             const float* w_d, int64_t w_bstride, float* field_d, void* stream);

/* leniax/core.py:245-319 get_state (v1 clip / v2 / simple); dt_d per world (dt_bstride 0: shared). */
int lx_state(const lx_rule* rule, const float* state_d, const float* field_d, int64_t B, const float* dt_d,
             int64_t dt_bstride, float* out_d, void* stream);

/* leniax/core.py:13-49 update with the FFT potential: the plain multi-pass step
 * (3 kernels, 2 half-spectrum passes through HBM).  field_d / potential_d may be NULL. */
int lx_update(const lx_rule* rule, const float* state_d, int64_t B, const void* K_half_d, int64_t k_bstride,
              const float* gf_d, int64_t gf_bstride, const float* w_d, int64_t w_bstride, const float* dt_d,
              int64_t dt_bstride, float* state_out_d, float* field_d, float* potential_d, void* work_d,
              void* stream);

/* lx_update that also writes the forward spectra of the input state, rfftn(state) as
 * [B][C][dims[:-1]][dims[-1]/2+1] complex64, to state_spec_d: the recompute step of the adjoint,
 * whose kernel-spectrum gradient needs them (lx_adjoint_potential). */
int lx_update_with_spectra(const lx_rule* rule, const float* state_d, int64_t B, const void* K_half_d,
                           int64_t k_bstride, const float* gf_d, int64_t gf_bstride, const float* w_d,
                           int64_t w_bstride, const float* dt_d, int64_t dt_bstride, float* state_out_d,
                           float* field_d, float* potential_d, void* state_spec_d, void* work_d, void* stream);

/* leniax/statistics.py:35-126 compute_stats (one step, B worlds) + the per-world carry
 * (shift, centroid, angle: leniax/runner.py:309-323) kept in carry_d (lx_carry_bytes() each).
 * stats_d [B][11 + C] floats (order: mass, mass_volume, mass_density, growth, growth_volume,
 * growth_density, mass_speed, mass_angle_speed, mass_growth_dist, inertia, potential_volume,
 * channel_mass[C]).  R and T are the config's values (leniax/statistics.py:22-24). */
size_t lx_carry_bytes(void);
int lx_carry_init(int64_t B, void* carry_d, void* stream);
int lx_stats(const lx_rule* rule, const float* cells_d, const float* field_d, const float* potential_d,
             int64_t B, float R, float T, void* carry_d, float* stats_d, void* stream);

/* One scan step of leniax/runner.py:295-334 (_scan_fn: update_fn, then compute_stats_fn on the
 * step's input state, field and potential): lx_update and lx_stats of the same step, with the
 * statistics summed by the inverse row pass while it holds the state, field and potential, so
 * that nothing is read back from HBM for them.  field_d / potential_d may be NULL
 * (run_scan_mem_optimized, leniax/runner.py:167-215, keeps neither).  work_d:
 * lx_workspace_size() bytes.  Results equal lx_update + lx_stats up to the order of the sums. */
int lx_step_stats(const lx_rule* rule, const float* state_d, int64_t B, const void* K_half_d, int64_t k_bstride,
                  const float* gf_d, int64_t gf_bstride, const float* w_d, int64_t w_bstride, const float* dt_d,
                  int64_t dt_bstride, float* state_out_d, float* field_d, float* potential_d, float R, float T,
                  void* carry_d, float* stats_d, void* work_d, void* stream);

/* leniax/statistics.py:134-205 check_heuristics over stats_d [T][B][11 + C] (lx_stats layout):
 * cont_d [T][B] (1.0 / 0.0) and n_d [B] = sum_t cont. */
int lx_check_heuristics(const float* stats_d, int32_t T, int64_t B, int32_t C, float* cont_d, float* n_d,
                        void* stream);

/* numpy.fft.rfft2 / irfft2 of n real 2-D arrays [n][H][W] <-> half spectra [n][H][W/2+1]
 * (the transforms leniax/kernels.py:140 and leniax/core.py:80-92 apply; irfft2 scales by 1/(H W)).
 * lx_irfft2 leaves X_d unchanged; work_d holds n * H * (W/2+1) complex64. */
int lx_rfft2(int32_t ndim, const int32_t* dims, int64_t n, const float* x_d, void* X_d, void* stream);
int lx_irfft2(int32_t ndim, const int32_t* dims, int64_t n, const void* X_d, float* x_d, void* work_d, void* stream);

/* Reverse mode of one update step: what jax.grad computes through leniax/core.py:13-49 inside
 * runner.make_pipeline_fn / make_gradient_fn (leniax/runner.py:363-444).  A step's adjoint is
 * lx_adjoint_pointwise followed by lx_adjoint_potential.
 *
 * lx_adjoint_pointwise: state_d, field_d, potential_d are the step's input state and the field and
 * potential lx_update wrote for it (one channel and one kernel: field_d may be NULL, the field is
 * then recomputed from the potential); lambda_d = dL/d(state out).  Writes the direct part of
 * dL/d(state in) to lambda_out_d (may alias lambda_d) and dL/d(potential) [B][K][dims] to dpot_d,
 * and ADDS dL/dgf, dL/dw, dL/ddt into grad_gf_d / grad_w_d / grad_dt_d (laid out and strided like
 * gf_d / w_d / dt_d; each may be NULL).  ste != 0: v1's clip passes gradients straight through, as
 * leniax/core.py:265-270 does; ste == 0: the derivative of the clip itself (1 inside (0, 1)).
 * dfield_d / dpot_in_d (NULL: zero) are cotangents of the step's field and potential outputs.
 * work_d: lx_adjoint_workspace_size() bytes, where the gradient sums are reduced in a fixed order
 * (the gradients are deterministic and do not depend on the batch a world is in). */
int lx_adjoint_workspace_size(const lx_rule* rule, int64_t B, size_t* bytes);
int lx_adjoint_pointwise(const lx_rule* rule, int64_t B, const float* state_d, const float* field_d,
                         const float* potential_d, const float* gf_d, int64_t gf_bstride, const float* w_d,
                         int64_t w_bstride, const float* dt_d, int64_t dt_bstride, const float* lambda_d,
                         const float* dfield_d, const float* dpot_in_d, int32_t ste, float* lambda_out_d,
                         float* dpot_d, float* grad_gf_d, float* grad_w_d, float* grad_dt_d, void* work_d,
                         void* stream);

/* The gradient of a whole run (leniax/runner.py:363-444: make_pipeline_fn + make_gradient_fn over
 * `steps` updates of leniax/core.py:13-49), with a checkpoint every ckpt_every steps and each
 * segment recomputed in the reverse pass.  lx_grad_run_forward writes the final states cells_d and
 * keeps the checkpoints in work_d; lx_grad_run_backward, given lambda_d = dL/d(final states) and
 * the same work_d, ADDS dL/dgf, dL/dw, dL/ddt (laid out and strided like gf / w / dt) and, when
 * grad_spec_d is set, sum over steps of rfftn(dL/dU_k) conj(rfftn(A_cin k)) per world ([B][K][half],
 * see lx_adjoint_potential for dL/dK and dL/d(real kernel)), and writes dL/d(initial states) to
 * grad_cells0_d.  Any gradient pointer may be NULL.  The backward reads what the forward left in
 * work_d (the checkpoints and the last segment's tensors): it follows the forward with the same
 * arguments, grad_spec_d NULL or not included, and does not read cells_d.  A forward with
 * forward_only = 1 only steps the worlds (no checkpoints, no segment kept). */
typedef struct lx_grad_run_args {
  const float* cells0_d;   /* [B][C][dims] initial states */
  const void* K_half_d;    /* lx_prepare_spectra layout */
  int64_t k_bstride;       /* complex elements between worlds' spectra, 0: shared */
  const float* gf_d;       /* [K][2] per world, stride gf_bstride floats */
  int64_t gf_bstride;
  const float* w_d;        /* [C][K] per world, stride w_bstride floats */
  int64_t w_bstride;
  const float* dt_d;       /* dt per world, stride dt_bstride */
  int64_t dt_bstride;
  int64_t B;               /* worlds */
  int32_t steps;           /* updates */
  int32_t ckpt_every;      /* steps between checkpoints */
  int32_t ste;             /* v1 clip: 1 straight-through (leniax/core.py:265-270), 0 its true derivative */
  float* cells_d;          /* [B][C][dims] final states (forward output) */
  const float* lambda_d;   /* [B][C][dims] dL/d(final states) (backward input) */
  float* grad_cells0_d;    /* [B][C][dims] dL/d(initial states) or NULL */
  float* grad_gf_d;        /* added into, or NULL */
  float* grad_w_d;
  float* grad_dt_d;
  void* grad_spec_d;       /* [B][K][half] complex64, added into, or NULL */
  void* work_d;            /* lx_grad_run_workspace_size() bytes: checkpoints + one segment */
  int32_t forward_only;    /* 1: no backward follows; the forward keeps nothing but writes cells_d */
} lx_grad_run_args;
int lx_grad_run_workspace_size(const lx_rule* rule, int64_t B, int32_t steps, int32_t ckpt_every, int32_t want_spec,
                               size_t* bytes);
int lx_grad_run_forward(const lx_rule* rule, const lx_grad_run_args* args, void* stream);
int lx_grad_run_backward(const lx_rule* rule, const lx_grad_run_args* args, void* stream);

/* Hermitian extension of n half spectra [n][H][W/2+1] to full spectra [n][H][W], times scale
 * (the full complex K layout of leniax/kernels.py:140 from a real array's half spectrum). */
int lx_half_to_full(int32_t ndim, const int32_t* dims, int64_t n, const void* half_d, void* full_d, float scale,
                    void* stream);

/* Adjoint of the FFT potential (leniax/core.py:52-102): lambda_d [B][C][dims] += for each kernel k
 * the correlation of dpot_d[b][k] with kernel k, i.e. irfft2(rfft2(dpot) * conj(K_half)).
 * With state_spec_d = rfft2 of the step's input state ([B][C][H][W/2+1]), also ADDS
 * rfft2(dpot[b][k]) * conj(state_spec[b][cin k]) into dspec_d [B][K][H][W/2+1] (per world, never
 * shared); irfft2 of its sum over steps is dL/d(real kernel), and its Hermitian extension divided
 * by H W is dL/dK for leniax's full complex K.  work_d: lx_workspace_size(rule, B) bytes. */
int lx_adjoint_potential(const lx_rule* rule, int64_t B, const void* K_half_d, int64_t k_bstride,
                         const float* dpot_d, const void* state_spec_d, void* dspec_d, float* lambda_d,
                         void* work_d, void* stream);

/* The reverse of one update step (lx_adjoint_pointwise then lx_adjoint_potential, with the same
 * arguments and results): lambda_d holds dL/d(state out) on entry and dL/d(state in) on return;
 * the gradient scalars and dspec are added into.  With one channel and one kernel the two are one
 * row pass (dL/dU is computed from the potential and transformed at once, never written; the field
 * is recomputed from the potential, so field_d and dpot_d are not read and may be NULL); otherwise
 * dpot_d ([B][K][dims]) and field_d are needed.  work_d:
 * lx_workspace_size() bytes, adj_work_d: lx_adjoint_workspace_size() bytes. */
int lx_adjoint_step(const lx_rule* rule, int64_t B, const float* state_d, const float* field_d,
                    const float* potential_d, const float* gf_d, int64_t gf_bstride, const float* w_d,
                    int64_t w_bstride, const float* dt_d, int64_t dt_bstride, const void* K_half_d, int64_t k_bstride,
                    int32_t ste, const void* state_spec_d, void* dspec_d, float* lambda_d, float* dpot_d,
                    float* grad_gf_d, float* grad_w_d, float* grad_dt_d, void* work_d, void* adj_work_d, void* stream);

/* Batched run (leniax/runner.py:167-215 run_scan_mem_optimized and :119-164 run_scan semantics):
 * every world steps up to max_steps, statistics and stop policy evaluated every step on the device,
 * worlds retiring at uneven times.  Two paths, chosen by the rule (lx_run_supported):
 *  - fused (2-D N x N, N in {32, 64, 128}, one channel, one kernel): one persistent launch steps
 *    world after world from a device queue, each world resident in registers + LDS;
 *  - multi-pass (every other rule lx_update takes, 2-D and 3-D): the step's HBM passes over a
 *    device-side list of live worlds; a control kernel after each step retires stopped worlds and
 *    refills their slots from the queue of worlds not yet started (at most `slots` in flight), and
 *    the step loop is replayed from HIP graphs of several steps.  This path waits on the stream
 *    between graph replays to read the live count (it returns once the last replay is enqueued).
 * The statistics rows after a world's last step are left as the caller put them. */
typedef struct lx_run_args {
  const float* cells0_d;   /* [B][C][dims] initial states */
  float* cells_d;          /* [B][C][dims] final states (may equal cells0_d) */
  const void* K_half_d;    /* lx_prepare_spectra layout */
  int64_t k_bstride;       /* complex elements between worlds' spectra, 0: shared */
  const float* gf_d;       /* [K][2] per world, stride gf_bstride floats */
  int64_t gf_bstride;
  const float* w_d;        /* [C][K] per world, stride w_bstride floats */
  int64_t w_bstride;
  const float* dt_d;       /* update dt = 1/T per world, stride dt_bstride */
  int64_t dt_bstride;
  int64_t B;               /* worlds */
  int32_t max_steps;
  int32_t stop_policy;     /* enum lx_stop */
  int32_t min_steps;       /* LX_STOP_LENIAX: never retire before this many steps */
  float R;                 /* statistics R and T (config values, leniax/statistics.py:22-24) */
  float T;
  float* stats_d;          /* [B][max_steps][11 + C] or NULL */
  int32_t* n_d;            /* [B] leniax N (leading steps whose heuristics held) or NULL */
  int32_t* steps_d;        /* [B] steps executed or NULL */
  float* cells_hist_d;     /* [max_steps][B][C][dims] state before each step, or NULL */
  float* field_hist_d;     /* [max_steps][B][C][dims] or NULL */
  float* potential_hist_d; /* [max_steps][B][K][dims] or NULL */
  void* scratch_d;         /* lx_run_scratch_size() bytes of device scratch */
  int32_t slots;           /* multi-pass path: worlds in flight at once (0: all B); ignored by the fused path */
} lx_run_args;

/* With an early-stop policy and more than 256 steps, lx_run is two launches: every world first
 * runs at most max_steps / 8 steps (a probe); the worlds still running then resume together from
 * their parked state and carry, so that the long-lived ones start at once instead of whenever a
 * compute unit frees up.  Results are those of one uninterrupted run. */
#define LX_RUN_FUSED 1
#define LX_RUN_PASSES 2
/* LX_RUN_FUSED or LX_RUN_PASSES: the path lx_run takes for this rule; 0: the rule is invalid or its
 * world size unsupported. */
int lx_run_supported(const lx_rule* rule);
/* Scratch bytes of lx_run for B worlds (slots as in lx_run_args). */
int lx_run_scratch_size(const lx_rule* rule, int64_t B, int32_t slots, size_t* bytes);
/* The fused path's scratch alone (= lx_run_scratch_size for a rule with LX_RUN_FUSED). */
size_t lx_run_scratch_bytes(int64_t B);
int lx_run(const lx_rule* rule, const lx_run_args* args, void* stream);

/* lx_run on the fused path with no device allocation and no workspace: args->scratch_d is ignored
 * and a fixed device array of the library (one per device, kept until the process ends) holds the
 * run's queues and parked carries.  No counterpart in leniax: it is for callers that cannot
 * allocate (code under HIP graph capture, serving loops, hosts that forbid hipMalloc).  At most
 * LX_RUN_STATIC_MAX_WORLDS worlds (LX_EINVAL beyond); rules without a fused kernel are LX_ENOTSUP.
 * The autotuner times nothing inside it (forced, recorded or default schedules only).
 * Every call on a device shares that device's array: calls that overlap on one device race. */
#define LX_RUN_STATIC_MAX_WORLDS 65536
int lx_run_static(const lx_rule* rule, const lx_run_args* args, void* stream);

/* Autotuner.  The passes have variants that compute the same thing in different schedules:
 *   "conv_staged"  the convolution pass at 256 points staged through LDS (1) or read in the lane
 *                  groups' layout (0);
 *   "line_xcd"     line passes' blocks dealt in XCD-contiguous ranges (1) or in launch order (0);
 *   "lines96_tpr"  3-D forward / inverse line passes at 96 points: lanes per line, 4 (24 points per
 *                  lane, transpose through LDS) or 8 (12 points, DPP network);
 *   "fused_probe"  fused early-stop run: first launch steps every world max_steps / value steps
 *                  (0: one launch);
 *   "graph_steps"  multi-pass lx_run: steps per captured graph.
 * LX_TUNE_OFF (the default, or LENIAX_AMD_TUNE=0): fixed rules.  LX_TUNE_ON: the first call of a
 * shape times the candidates on its own inputs (outside stream capture) and keeps the fastest, in
 * memory and in the cache file (LENIAX_AMD_TUNE_CACHE or lx_tune_cache).  LX_TUNE_CACHED: decisions
 * from memory / the cache file, fixed rules otherwise.  lx_tune_force pins a knob for every shape
 * (value < 0 unpins).  lx_autotune returns the previous mode; lx_tune_report writes "knob key value"
 * lines (forced knobs first) and returns the text's length.  No counterpart in leniax, whose
 * schedules are XLA's (leniax/core.py:76-87). */
#define LX_TUNE_OFF 0
#define LX_TUNE_ON 1
#define LX_TUNE_CACHED 2
int lx_autotune(int32_t mode);
int lx_tune_cache(const char* path);
int lx_tune_force(const char* knob, int32_t value);
int lx_tune_report(char* buf, size_t n);

#ifdef __cplusplus
}
#endif
#endif
