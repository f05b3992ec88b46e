// The gradient of a whole run for C callers: what jax.grad gives leniax through
// runner.make_pipeline_fn / make_gradient_fn (leniax/runner.py:363-444) over `steps` updates of
// leniax/core.py:13-49, with checkpoints every `ckpt_every` steps and recomputation.
//
// lx_grad_run_forward steps the worlds from their initial states, keeping every ckpt_every-th
// state in the workspace, and writes the final states; the last segment's steps also keep their
// input states, potentials (fields, spectra) in the segment buffers, which the reverse pass reads
// first.  With forward_only it keeps nothing (no reverse pass follows).  lx_grad_run_backward takes dL/d(final state) and walks the segments from the last: it
// recomputes each earlier segment from its checkpoint, keeping each step's input state, potential,
// field (with several channels or kernels; one channel and one kernel recompute it from the
// potential) and the state's spectra (when the kernel gradient is asked for), then runs the step's
// reverse (lx_adjoint_step) from the segment's last step to its first.
// Memory: the checkpoints plus one segment, instead of every step's tensors.  The kernels and their
// order are those of the per-step entry points.
#include <algorithm>

#include "lx_internal.h"

namespace {

struct GradLayout {
  size_t ckpt, seg_A, seg_G, seg_U, seg_F, lam, dpot, step_work, adj_work, total;
  int64_t state_elems, pot_elems, spec_elems;  // per batch: B C n, B K n, B C half
  int nckpt;
};

int64_t cells_of_rule(const lx_rule* r) {
  int64_t n = 1;
  for (int i = 0; i < r->ndim; ++i) n *= r->dims[i];
  return n;
}

int grad_layout(const lx_rule* r, int64_t B, int32_t steps, int32_t c, int32_t want_spec, GradLayout& l) {
  if (steps < 1 || c < 1) return lx_fail(LX_EINVAL, "steps and ckpt_every must be positive");
  auto up = [](size_t v) { return (v + 255) / 256 * 256; };
  const int64_t n = cells_of_rule(r);
  int64_t rows = 1;
  for (int i = 0; i < r->ndim - 1; ++i) rows *= r->dims[i];
  const int64_t half = rows * (r->dims[r->ndim - 1] / 2 + 1);
  const int seg = std::min(c, steps);
  l.state_elems = B * r->C * n;
  l.pot_elems = B * r->K * n;
  l.spec_elems = want_spec ? B * r->C * half : 0;
  l.nckpt = (steps + c - 1) / c;
  size_t step_work = 0, adj_work = 0;
  if (int rc = lx_workspace_size(r, B, &step_work)) return rc;
  if (int rc = lx_adjoint_workspace_size(r, B, &adj_work)) return rc;
  size_t o = 0;
  l.ckpt = o;      o = up(o + size_t(l.nckpt) * l.state_elems * sizeof(float));
  l.seg_A = o;     o = up(o + size_t(seg + 1) * l.state_elems * sizeof(float));  // + the segment's output
  const bool keepG = !(r->C == 1 && r->K == 1);  // one channel, one kernel: the field is recomputed
  l.seg_G = o;     o = up(o + (keepG ? size_t(seg) * l.state_elems * sizeof(float) : 0));
  l.seg_U = o;     o = up(o + size_t(seg) * l.pot_elems * sizeof(float));
  l.seg_F = o;     o = up(o + size_t(seg) * l.spec_elems * sizeof(float2));
  l.lam = o;       o = up(o + size_t(l.state_elems) * sizeof(float));
  l.dpot = o;      o = up(o + size_t(l.pot_elems) * sizeof(float));
  l.step_work = o; o = up(o + step_work);
  l.adj_work = o;  o = up(o + adj_work);
  l.total = o;
  return LX_OK;
}

int check_args(const lx_rule* rule, const lx_grad_run_args* a, bool forward) {
  int rc = lx_check_rule(rule);
  if (rc || (rc = lx_no_quadrant(rule, "lx_grad_run"))) return rc;
  if (!a) return lx_fail(LX_EINVAL, "args is NULL");
  if (a->B < 0 || a->steps < 1 || a->ckpt_every < 1) return lx_fail(LX_EINVAL, "B >= 0, steps and ckpt_every >= 1");
  if (a->B > 0 && (!a->cells0_d || !a->K_half_d || !a->gf_d || !a->w_d || !a->dt_d || (forward && !a->cells_d) ||
                   !a->work_d))
    return lx_fail(LX_EINVAL, "NULL device pointer in lx_grad_run_args");
  return LX_OK;
}

float* fptr(void* base, size_t off) { return reinterpret_cast<float*>(static_cast<char*>(base) + off); }

int copy(float* dst, const float* src, int64_t elems, hipStream_t st) {
  if (hipMemcpyAsync(dst, src, size_t(elems) * sizeof(float), hipMemcpyDeviceToDevice, st) != hipSuccess)
    return lx_fail(LX_EDEVICE, "device copy failed");
  return LX_OK;
}

// one step of a segment with its potential (field, spectra) kept in the segment buffers
int segment_step(const lx_rule* rule, const lx_grad_run_args* a, const GradLayout& l, int i, const float* Ai,
                 float* out, void* work, void* stream) {
  const bool spec = a->grad_spec_d != nullptr;
  const bool keepG = !(rule->C == 1 && rule->K == 1);
  float* Gi = keepG ? fptr(a->work_d, l.seg_G) + int64_t(i) * l.state_elems : nullptr;
  float* Ui = fptr(a->work_d, l.seg_U) + int64_t(i) * l.pot_elems;
  if (spec)
    return lx_update_with_spectra(rule, Ai, a->B, a->K_half_d, a->k_bstride, a->gf_d, a->gf_bstride, a->w_d,
                                  a->w_bstride, a->dt_d, a->dt_bstride, out, Gi, Ui,
                                  reinterpret_cast<float2*>(fptr(a->work_d, l.seg_F)) + int64_t(i) * l.spec_elems,
                                  work, stream);
  return lx_update(rule, Ai, a->B, a->K_half_d, a->k_bstride, a->gf_d, a->gf_bstride, a->w_d, a->w_bstride,
                   a->dt_d, a->dt_bstride, out, Gi, Ui, work, stream);
}

}  // namespace

extern "C" int lx_grad_run_workspace_size(const lx_rule* rule, int64_t B, int32_t steps, int32_t ckpt_every,
                                          int32_t want_spec, size_t* bytes) {
  int rc = lx_check_rule(rule);
  if (rc) return rc;
  if (!bytes || B < 0) return lx_fail(LX_EINVAL, "bad arguments");
  GradLayout l;
  if ((rc = grad_layout(rule, B, steps, ckpt_every, want_spec, l))) return rc;
  *bytes = l.total;
  return LX_OK;
}

extern "C" int lx_grad_run_forward(const lx_rule* rule, const lx_grad_run_args* a, void* stream) {
  int rc = check_args(rule, a, true);
  if (rc || a->B == 0) return rc;
  GradLayout l;
  if ((rc = grad_layout(rule, a->B, a->steps, a->ckpt_every, a->grad_spec_d != nullptr, l))) return rc;
  hipStream_t st = (hipStream_t)stream;
  float* ck = fptr(a->work_d, l.ckpt);
  float* sA = fptr(a->work_d, l.seg_A);
  void* work = fptr(a->work_d, l.step_work);
  const bool only = a->forward_only != 0;
  if (!only && (rc = copy(ck, a->cells0_d, l.state_elems, st))) return rc;
  // forward only: every step here, the last into cells_d; else the segments before the last, their
  // states at multiples of ckpt_every going to the checkpoints
  const int c = a->ckpt_every, last = only ? a->steps : (l.nckpt - 1) * c;
  const float* cur = only ? a->cells0_d : ck;
  for (int t = 0; t < last; ++t) {
    float* next = t + 1 == a->steps ? a->cells_d
                  : (t + 1) % c == 0 ? ck + int64_t((t + 1) / c) * l.state_elems
                                     : sA + int64_t(t & 1) * l.state_elems;
    if ((rc = lx_update(rule, cur, a->B, a->K_half_d, a->k_bstride, a->gf_d, a->gf_bstride, a->w_d, a->w_bstride,
                        a->dt_d, a->dt_bstride, next, nullptr, nullptr, work, stream)))
      return rc;
    cur = next;
  }
  if (only) return LX_OK;
  // the last segment, its tensors kept for the reverse pass
  if ((rc = copy(sA, cur, l.state_elems, st))) return rc;
  const int L = a->steps - last;
  for (int i = 0; i < L; ++i) {
    float* out = i + 1 == L ? a->cells_d : sA + int64_t(i + 1) * l.state_elems;
    if ((rc = segment_step(rule, a, l, i, sA + int64_t(i) * l.state_elems, out, work, stream))) return rc;
  }
  return LX_OK;
}

extern "C" int lx_grad_run_backward(const lx_rule* rule, const lx_grad_run_args* a, void* stream) {
  int rc = check_args(rule, a, false);
  if (rc || a->B == 0) return rc;
  if (!a->lambda_d) return lx_fail(LX_EINVAL, "lx_grad_run_backward: lambda_d is NULL");
  GradLayout l;
  const bool spec = a->grad_spec_d != nullptr;
  if ((rc = grad_layout(rule, a->B, a->steps, a->ckpt_every, spec, l))) return rc;
  hipStream_t st = (hipStream_t)stream;
  float* ck = fptr(a->work_d, l.ckpt);
  float* sA = fptr(a->work_d, l.seg_A);
  float* sG = fptr(a->work_d, l.seg_G);
  float* sU = fptr(a->work_d, l.seg_U);
  float2* sF = reinterpret_cast<float2*>(fptr(a->work_d, l.seg_F));
  float* lam = fptr(a->work_d, l.lam);
  float* dpot = fptr(a->work_d, l.dpot);
  void* work = fptr(a->work_d, l.step_work);
  void* adj_work = fptr(a->work_d, l.adj_work);
  if ((rc = copy(lam, a->lambda_d, l.state_elems, st))) return rc;
  const int c = a->ckpt_every;
  const bool keepG = !(rule->C == 1 && rule->K == 1);
  // one channel and one kernel: the steps' reverses chained through fused row passes (each step's
  // last row pass shares a kernel with the first of the step before it, inside a segment), the
  // gradient scalars accumulated per block over the steps and summed once at the end
  const bool chain = !keepG && a->B <= 65535;
  LxAdj11 ac;
  if (chain && (rc = lx_adj11_init(rule, a->B, a->gf_d, a->gf_bstride, a->w_d, a->w_bstride, a->dt_d,
                                   a->dt_bstride, a->K_half_d, a->k_bstride, a->ste, lam, a->grad_gf_d,
                                   a->grad_w_d, a->grad_dt_d, work, adj_work, stream, &ac)))
    return rc;
  for (int seg = l.nckpt - 1; seg >= 0; --seg) {
    const int t0 = seg * c, L = std::min(c, a->steps - t0);
    // recompute the segment (the last one the forward kept): input state, potential (field, spectra)
    // of each of its steps
    if (seg < l.nckpt - 1) {
      if ((rc = copy(sA, ck + int64_t(seg) * l.state_elems, l.state_elems, st))) return rc;
      for (int i = 0; i < L; ++i)
        if ((rc = segment_step(rule, a, l, i, sA + int64_t(i) * l.state_elems, sA + int64_t(i + 1) * l.state_elems,
                               work, stream)))
          return rc;
    }
    if (chain) {
      for (int i = L - 1; i >= 0; --i) {
        const float* Ui = sU + int64_t(i) * l.pot_elems;
        if (i == L - 1 &&
            (rc = lx_adj11_rows(ac, Ui, sA + int64_t(i) * l.state_elems, false, seg < l.nckpt - 1)))
          return rc;
        if ((rc = lx_adj11_lines(ac, spec ? sF + int64_t(i) * l.spec_elems : nullptr, a->grad_spec_d))) return rc;
        if (i > 0) {
          if ((rc = lx_adj11_rows(ac, sU + int64_t(i - 1) * l.pot_elems, sA + int64_t(i - 1) * l.state_elems, true,
                                  true)))
            return rc;
        } else if ((rc = lx_adj11_last(ac, seg == 0))) {
          return rc;
        }
      }
      continue;
    }
    // reverse, last step first: lambda updated in place
    for (int i = L - 1; i >= 0; --i) {
      const float* Ai = sA + int64_t(i) * l.state_elems;
      if ((rc = lx_adjoint_step(rule, a->B, Ai, keepG ? sG + int64_t(i) * l.state_elems : nullptr,
                                sU + int64_t(i) * l.pot_elems, a->gf_d, a->gf_bstride, a->w_d, a->w_bstride, a->dt_d,
                                a->dt_bstride, a->K_half_d, a->k_bstride, a->ste,
                                spec ? sF + int64_t(i) * l.spec_elems : nullptr, a->grad_spec_d, lam, dpot,
                                a->grad_gf_d, a->grad_w_d, a->grad_dt_d, work, adj_work, stream)))
        return rc;
    }
  }
  if (a->grad_cells0_d) return copy(a->grad_cells0_d, lam, l.state_elems, st);
  return LX_OK;
}
