// placeholder, replaced by the MFMA implementation
#include "api_internal.h"
extern "C" {
struct gpsm_mlp { int dummy; };
int gpsm_mlp_create(gpsm_ctx*, uint64_t, gpsm_mlp** out) { if (out) *out = nullptr; return GPSM_EINVAL; }
void gpsm_mlp_free(gpsm_mlp*) {}
uint64_t gpsm_mlp_num_params(void) { return 0; }
int gpsm_mlp_get_params(gpsm_mlp*, float*) { return GPSM_EINVAL; }
int gpsm_mlp_set_params(gpsm_mlp*, const float*) { return GPSM_EINVAL; }
int gpsm_mlp_forward(gpsm_mlp*, const float*, uint32_t, float*) { return GPSM_EINVAL; }
int gpsm_mlp_grad(gpsm_mlp*, const float*, const float*, uint32_t, float*, float*) { return GPSM_EINVAL; }
int gpsm_mlp_train_step(gpsm_mlp*, const float*, const float*, uint32_t, float, float*) { return GPSM_EINVAL; }
int gpsm_set_planner(gpsm_ctx*, gpsm_mlp*) { return GPSM_EINVAL; }
int gpsm_plan_features(uint32_t, const gpsm_query*, const uint32_t*, const uint32_t*, const uint32_t*, uint32_t, uint32_t, uint64_t, const uint8_t*, float, float*) { return GPSM_EINVAL; }
}
