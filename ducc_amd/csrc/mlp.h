// mlp.h -- the chooser network: sizes and kernel launchers (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#define MLP_IN 6
#define MLP_HID 64
#define MLP_OUT 3
#define MLP_NPARAM                                                           \
  (MLP_HID * MLP_IN + MLP_HID + 2 * (MLP_HID * MLP_HID + MLP_HID) +         \
   MLP_OUT * MLP_HID + MLP_OUT)

namespace dmx {
// act: B * mlp_act_floats(1) floats (pre-activations + outputs), may be null
void mlp_forward(const float *P, const float *X, int B, float *act, float *Y,
                 hipStream_t s);
// dz must hold mlp_act_floats(B) + B*MLP_OUT floats; lossb B*MLP_OUT floats
void mlp_grad(const float *P, const float *X, const float *T,
              const float *wout, int B, float *act, float *dz, float *lossb,
              float *G, float *loss, hipStream_t s);
void mlp_adam_step(float *P, const float *G, float *m, float *v, float lr,
                   int t, float wd, hipStream_t s);
size_t mlp_act_floats(int B);
}  // namespace dmx
