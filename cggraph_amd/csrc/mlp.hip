// Learned scheduler kernels (filled in below).
#include "engine.hpp"

extern "C" {
int cg_mlp_train(const char*, const char*, int, float, uint64_t, float*) {
  return cg::set_error(CG_ERR_UNSUPPORTED, "MLP training is not built yet");
}
int cg_mlp_forward(const float*, const float*, int, float*) {
  return cg::set_error(CG_ERR_UNSUPPORTED, "MLP forward is not built yet");
}
int cg_mlp_loss_grad(const float*, const float*, const int*, int, float*, float*) {
  return cg::set_error(CG_ERR_UNSUPPORTED, "MLP gradients are not built yet");
}
}
