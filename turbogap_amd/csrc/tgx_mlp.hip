// Surrogate and MLP training entry points (filled in below).
#include "tgx_internal.h"

int tgx_surrogate_run(tgx_ctx* ctx, tgx_model_t* m) {
  (void)ctx; (void)m;
  tgx_set_error("surrogate path not built yet");
  return TGX_ERR_UNSUPPORTED;
}

#define TGX_NOT_YET(sig)                                           \
  extern "C" int sig {                                             \
    tgx_set_error("%s: not available in this build", __func__);    \
    return TGX_ERR_UNSUPPORTED;                                    \
  }
TGX_NOT_YET(tgx_surrogate_set(tgx_ctx*, int, int, int, const float*, const float*, const float*, const float*,
                              const float*, const float*, double, double))
TGX_NOT_YET(tgx_surrogate_gate(tgx_ctx*, int, double))
TGX_NOT_YET(tgx_surrogate_stats(tgx_ctx*, double*, double*))
TGX_NOT_YET(tgx_mlp_create(tgx_ctx*, int, int, int, uint64_t, tgx_mlp**))
TGX_NOT_YET(tgx_mlp_destroy(tgx_mlp*))
TGX_NOT_YET(tgx_mlp_train_step(tgx_mlp*, int, const float*, const float*, const float*, float, float, float, double*,
                               double*))
TGX_NOT_YET(tgx_mlp_forward(tgx_mlp*, int, const float*, float*, float*))
TGX_NOT_YET(tgx_mlp_get_weights(tgx_mlp*, float*, float*, float*, float*, float*, float*))
TGX_NOT_YET(tgx_mlp_set_weights(tgx_mlp*, const float*, const float*, const float*, const float*, const float*,
                                const float*))
TGX_NOT_YET(tgx_model_training_view(tgx_ctx*, int, int*, const float**, const float**, const double**, const int**))
