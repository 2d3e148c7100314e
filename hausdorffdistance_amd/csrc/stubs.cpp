// Temporary: variants not yet built return HD_EUNSUPPORTED.
#include "hd_internal.h"

namespace hd {
int RunHybrid(hd_context* ctx, const DirectedArgs& a, hd_directed_result* o) {
  return RunEarlyBreak(ctx, a, o);
}
int LearnedOrder(hd_context*, const DirectedArgs&, const uint32_t**) {
  return Fail(HD_EUNSUPPORTED, "learned order not built yet");
}
}  // namespace hd

extern "C" {
int hd_batch_directed_device(hd_context*, int, const float*, const uint64_t*,
                             const float*, const uint64_t*, uint32_t, uint32_t,
                             hd_directed_result*) {
  return hd::Fail(HD_EUNSUPPORTED, "batch not built yet");
}
void hd_mlp_train_config_init(hd_mlp_train_config* c) {}
int hd_mlp_train_device(hd_context*, int, const float*, uint64_t, const float*,
                        uint64_t, const hd_mlp_train_config*, hd_mlp**, float*) {
  return hd::Fail(HD_EUNSUPPORTED, "mlp not built yet");
}
int hd_mlp_predict_device(hd_context*, const hd_mlp*, const float*, uint64_t,
                          float*) {
  return hd::Fail(HD_EUNSUPPORTED, "mlp not built yet");
}
void hd_mlp_destroy(hd_mlp*) {}
int hd_mlp_n_params(int n_dims) { return 64 * n_dims + 64 + 64 * 64 + 64 + 64 + 1; }
int hd_mlp_get_params(const hd_mlp*, float*, float*, float*) {
  return hd::Fail(HD_EUNSUPPORTED, "mlp not built yet");
}
int hd_mlp_loss_grad_device(hd_context*, int, const float*, const float*,
                            const float*, uint32_t, float*, float*) {
  return hd::Fail(HD_EUNSUPPORTED, "mlp not built yet");
}
}
