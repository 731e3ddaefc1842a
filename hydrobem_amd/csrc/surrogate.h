// The learned surrogate of the finite-depth evanescent remainder (surrogate.hip): sizes, the
// parameter layout, the input map, the device copy of the packed weights.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace hbg {

constexpr int NN_IN = 4, NN_H = 128, NN_OUT = 4;
// FP32 parameters, output-major weights: W1 [128][4], b1, W2 [128][128], b2, W3 [128][128], b3,
// W4 [4][128], b4
constexpr int NN_OFF_W1 = 0;
constexpr int NN_OFF_B1 = NN_OFF_W1 + NN_H * NN_IN;
constexpr int NN_OFF_W2 = NN_OFF_B1 + NN_H;
constexpr int NN_OFF_B2 = NN_OFF_W2 + NN_H * NN_H;
constexpr int NN_OFF_W3 = NN_OFF_B2 + NN_H;
constexpr int NN_OFF_B3 = NN_OFF_W3 + NN_H * NN_H;
constexpr int NN_OFF_W4 = NN_OFF_B3 + NN_H;
constexpr int NN_OFF_B4 = NN_OFF_W4 + NN_OUT * NN_H;
constexpr int NN_PARAMS = NN_OFF_B4 + NN_OUT;

// the training domain: R/h in [1e-2, 20], nu h in [0.05, 20] (log-uniform), z/h, zeta/h in [-1, 0]
constexpr double NN_LNR_LO = -4.605170185988091;   // ln 1e-2
constexpr double NN_LNR_HI = 2.995732273553991;    // ln 20
constexpr double NN_LNNU_LO = -2.995732273553991;  // ln 0.05
constexpr double NN_LNNU_HI = 2.995732273553991;   // ln 20
// the assembly uses the network for R/h >= NN_RMIN only (the series/tiles below)
constexpr double NN_RMIN = 0.1;

// the network's input of (R/h, z/h, zeta/h, nu h)
__host__ __device__ inline float4 nn_input(double r, double z, double zeta, double nuh) {
  return make_float4((float)(2.0 * (log(r) - NN_LNR_LO) / (NN_LNR_HI - NN_LNR_LO) - 1.0), (float)(2.0 * z + 1.0),
                     (float)(2.0 * zeta + 1.0),
                     (float)(2.0 * (log(nuh) - NN_LNNU_LO) / (NN_LNNU_HI - NN_LNNU_LO) - 1.0));
}

// The network's outputs are (h T / s0, h^2 T_R / s1, h^2 T_zeta / s1, h^2 T_z / s1) with
// s0 = 1 + nu h and s1 = (1 + nu h)^2 / min(1, R/h): the remainder grows like nu h and its
// derivatives like (nu h)^2 and like 1 / R near the axis.
__host__ __device__ inline void nn_scales(double r, double nuh, double& s0, double& s1) {
  s0 = 1.0 + nuh;
  s1 = s0 * s0 / (r < 1.0 ? r : 1.0);
}

// the packed BF16 weights and FP32 biases of the inference kernels, and the FP32 parameters
struct NnDevice {
  void* pk = nullptr;
  float* bias = nullptr;
  std::vector<float> params;
  bool ready = false;
  ~NnDevice();
  void release();
  void upload(const float* params);
};

}  // namespace hbg
