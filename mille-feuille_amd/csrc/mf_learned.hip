// mf_learned.hip -- learned tile-precision policy and batched solves.
// (Entry points are being built; until then they report MF_ERR_UNSUPPORTED.)
#include "mf_device.h"

using namespace mf;

extern "C" mf_status mf_mlp_forward(const float *, const float *, int, float *, void *) {
  return fail(MF_ERR_UNSUPPORTED, "mf_mlp_forward: not built yet");
}
extern "C" mf_status mf_mlp_grad(const float *, const float *, const int *, int, float *, float *,
                                 void *) {
  return fail(MF_ERR_UNSUPPORTED, "mf_mlp_grad: not built yet");
}
extern "C" mf_status mf_mlp_train(float *, const float *, const int *, int, int, int, float,
                                  unsigned, float *, void *) {
  return fail(MF_ERR_UNSUPPORTED, "mf_mlp_train: not built yet");
}
extern "C" mf_status mf_policy_set_mlp(mf_tiled *, const float *) {
  return fail(MF_ERR_UNSUPPORTED, "mf_policy_set_mlp: not built yet");
}
extern "C" mf_status mf_policy_features(mf_tiled *, const double *, double, int, float *, void *) {
  return fail(MF_ERR_UNSUPPORTED, "mf_policy_features: not built yet");
}
extern "C" mf_status mf_batch_cg_solve(mf_tiled **, int, int, const double **, double **,
                                       const mf_solve_opts *, int *, double *, double *, double *) {
  return fail(MF_ERR_UNSUPPORTED, "mf_batch_cg_solve: not built yet");
}
