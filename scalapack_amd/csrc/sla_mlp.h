// sla_mlp.h -- the learned precision selector: its network (mlp.hip) and its
// features (aux.hip).
#pragma once

#include "sla_handle.h"

namespace sla {

constexpr int MLP_IN = 12, MLP_HIDDEN = 64, MLP_OUT = 7;
// W1, b1, W2, b2, W3, b3, then the input normalisation mu, sigma
constexpr int MLP_NPARAM = MLP_HIDDEN * MLP_IN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN +
                           MLP_OUT * MLP_HIDDEN + MLP_OUT + 2 * MLP_IN;

// Mean cost-weighted softmax cross-entropy of B samples (X [B][12] raw
// features, labels y in 0..MLP_OUT-1, cost [MLP_OUT] or null) and its gradient with
// respect to the trained parameters (the first MLP_NPARAM - 24 floats of dG).
float mlp_grad_device(Handle &h, int B, const float *dX, const int *dy, const float *dcost,
                      const float *dP, float *dG);
// Full-batch Adam for `epochs` steps, parameters updated in place;
// loss_history (host, epochs floats) may be null.
void mlp_train_device(Handle &h, int B, const float *dX, const int *dy, const float *dcost,
                      float *dP, int epochs, float lr, float *loss_history);
// Class (argmax) and probabilities [B][MLP_OUT] (either may be null).
void mlp_predict_device(Handle &h, int B, const float *dX, const float *dP, int *dpred,
                        float *dprob);

// The 12 selector features of an n x n FP64 matrix (host output).
void selector_features_device(Handle &h, int n, const double *dA, int lda, float *feat);

}  // namespace sla
