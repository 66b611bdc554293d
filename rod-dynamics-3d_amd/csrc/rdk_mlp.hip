// rdk_mlp.hip -- learned PSOR warm start (placeholder until the network
// kernels land): warm_mode 2 reports that no network is loaded.
#include "rdk_internal.hpp"

namespace rdk {

int mlp_warm_start(rdk_ctx* ctx) {
  ctx->err = "learned warm start: no network loaded";
  return RDK_ERR_STATE;
}

void mlp_free(rdk_ctx* ctx) { ctx->mlp = nullptr; }

}  // namespace rdk
