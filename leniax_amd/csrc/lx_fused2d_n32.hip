// Fused persistent run, 32 x 32 worlds (all growth functions, with and without history output).
#include "lx_fused2d.h"

int lx_fused2d_launch_32(const lx_rule* rule, const lx_run_args* a, hipStream_t stream) {
  return lx_fused::launch<32>(rule, a, stream);
}
