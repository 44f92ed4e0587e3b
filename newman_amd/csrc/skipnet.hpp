// skipnet.hpp - the learned series-skip predictor (see skipnet.hip).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/newman_amd.h"

namespace nm {

class SkipNet {
 public:
  static constexpr int kIn = 132;     // features per tile
  static constexpr int kHidden = 64;  // two hidden layers
  static constexpr int kParams = kHidden * kIn + kHidden + kHidden * kHidden + kHidden + kHidden + 1;  // 12737

  struct TrainOpts {
    int epochs = 200;
    int batch = 256;
    float lr = 3e-3f;
    uint64_t seed = 1;
  };

  SkipNet();
  ~SkipNet();
  SkipNet(const SkipNet&) = delete;
  SkipNet& operator=(const SkipNet&) = delete;

  void init(uint64_t seed);
  // Standardises the features on this data, initialises from opts.seed and runs
  // Adam on the GPU; curve gets the first batch's loss of every epoch.
  int train(const float* raw, const float* t, size_t n, const TrainOpts& o, std::vector<float>* curve);
  // y[i] for raw (unnormalised) features, on the GPU
  int predict(const float* raw, size_t n, float* y);
  // mean-squared-error loss and its gradient for normalised inputs, on the GPU
  int gradient(const float* xn, const float* t, size_t n, float* grad, float* loss);
  void normalise(const float* raw, size_t n, std::vector<float>& out) const;
  void fit_normaliser(const float* raw, size_t n);
  const char* last_error() const { return err_.c_str(); }

  std::vector<float> params;        // W1, b1, W2, b2, W3, b3
  std::vector<float> mean, scale;   // feature standardisation
  bool trained = false;
  uint64_t version = 0;  // bumped whenever the parameters change

 private:
  int reserve(size_t nx, size_t nb);
  int batch_grad(int n);
  void release();
  std::string err_;
  float *dP = nullptr, *dG = nullptr, *dM = nullptr, *dV = nullptr, *dX = nullptr, *dT = nullptr;
  float *dXB = nullptr, *dTB = nullptr, *dH1 = nullptr, *dH2 = nullptr, *dY = nullptr, *dDY = nullptr;
  float *dD1 = nullptr, *dD2 = nullptr, *dLoss = nullptr;
  int* dIdx = nullptr;
  size_t capX = 0, capB = 0;
};

}  // namespace nm
