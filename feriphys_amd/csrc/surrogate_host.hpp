// Host object of the learned surrogate network (fp_surrogate): master weights and optimiser state
// on the device, the device training set and the training workspace; plus the packing of the
// inference kernel's LDS image.
#pragma once

#include <vector>

#include "host_util.hpp"
#include "surrogate.hpp"

struct fp_surrogate {
  fp_surrogate_config cfg{};
  hipStream_t stream = nullptr;
  fph::DBuf<float> p, g, m, v;  // f32 master weights, gradient, Adam moments (kParams each)
  fph::DBuf<__bf16> wb;         // bf16 copy of the weights read by the training GEMMs
  uint64_t adam_t = 0;
  // training set: features (bf16 [count][72]) and targets (f32 [count][3], output units)
  fph::DBuf<__bf16> X;
  fph::DBuf<float> T;
  uint64_t count = 0;
  // workspace for `rows` samples
  int rows = 0;
  fph::DBuf<__bf16> Xb, H1, H2, H3, dY, dZ;
  fph::DBuf<float> Tb, Z1, Z2, Z3, Y, dH, partial, loss;

  ~fp_surrogate() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace fph {

// Forward + MSE + backward for the first `rows` samples in s->Xb / s->Tb; the gradient lands in
// s->g and the batch loss in s->loss.p[loss_slot].
int surrogate_fwd_bwd(fp_surrogate* s, int rows, int loss_slot);
// The inference kernel's LDS image (kImgBytes) from the f32 parameter vector.
std::vector<uint8_t> pack_image(const std::vector<float>& params);

}  // namespace fph
