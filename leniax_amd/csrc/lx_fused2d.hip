// Dispatch of the fused persistent 2-D run (kernel: lx_fused2d.h, one translation unit per size).
#include "lx_internal.h"

int lx_fused2d_launch_32(const lx_rule* rule, const lx_run_args* a, hipStream_t stream);
int lx_fused2d_launch_64(const lx_rule* rule, const lx_run_args* a, hipStream_t stream);
int lx_fused2d_launch_128(const lx_rule* rule, const lx_run_args* a, hipStream_t stream);

int lx_fused2d_supported(const lx_rule* rule) {
  if (rule->ndim != 2 || rule->C != 1 || rule->K != 1 || rule->k_quadrant) return 0;
  if (rule->dims[0] != rule->dims[1]) return 0;
  const int n = rule->dims[0];
  return n == 32 || n == 64 || n == 128;
}

int lx_fused2d_launch(const lx_rule* rule, const lx_run_args* a, hipStream_t stream) {
  switch (rule->dims[0]) {
    case 32: return lx_fused2d_launch_32(rule, a, stream);
    case 64: return lx_fused2d_launch_64(rule, a, stream);
    case 128: return lx_fused2d_launch_128(rule, a, stream);
    default: return lx_fail(LX_ENOTSUP, "fused run: unsupported world size %d", rule->dims[0]);
  }
}
