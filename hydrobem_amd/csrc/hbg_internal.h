// Internal declarations of the MI355X library (gfx950, HIP): device data layouts and the
// entry points shared by the translation units. Not part of the C ABI (include/hydrobem_amd.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace hbg {

constexpr double PI = 3.14159265358979323846;
constexpr double GAMMA = 0.5772156649015329;  // Euler's constant
constexpr double LOG2 = 0.69314718055994530942;

// ------------------------------------------------------------------ panels in HBM
// One wetted panel after the projection rule, as every kernel reads it (an array of structures:
// the statics and assembly kernels read one panel per wavefront-uniform index, so the 448-byte
// record is a broadcast load). Field order is fixed; doubles only.
struct PanelD {
  double c[3];      // area centroid
  double n[3];      // unit normal, into the fluid
  double area;
  double radius;    // max |v - c| over the distinct vertices
  double v[4][3];   // distinct projected vertices, counter-clockwise about n; v[3] = v[2] for a triangle
  double elen[4];   // edge lengths
  double em[4][3];  // outward in-plane unit normals of the edges
  double q4[4][3];  // the four projected corners as given (bilinear map of the 3 x 3 far rule)
  double nv;        // 3 or 4 distinct vertices
  double top;       // largest vertex z
  double interior;  // 1: a panel of the interior free surface, 2: a lid panel (both in z = 0, normal +z),
                    // 0: a body panel
  double lid_eps;   // a lid panel's damping factor eps (kappa = w^2 / (g + i w eps)), else 0
  double pad[4];
};
static_assert(sizeof(PanelD) == 56 * sizeof(double), "PanelD layout");

// One static entry: Rankine source and dipole of 1/r + 1/r' and the two log corrections.
struct alignas(32) StaticD {
  double rs, rd, dl, dd;
};

// ------------------------------------------------------------------ wave-kernel tables
// Infinite depth: unit tiles of [0, XT) x [0, YT), monomial coefficients in the tile variable
// t = 2 (x - i) - 1 in [-1, 1].
constexpr int XT = 48;
constexpr int YT = 48;
constexpr int D2 = 10;           // degree of the 2D tiles of V and V_X
constexpr int M2 = D2 + 1;
constexpr int D1 = 14;           // degree of the 1D tiles of J0, J1, Yt, Yt'
constexpr int M1 = D1 + 1;
constexpr int TILE2 = 2 * M2 * M2;  // doubles per 2D tile: [function][y power][x power]
constexpr int TILE1 = 4 * M1;       // doubles per 1D tile: [power][function]

struct WaveTables {
  const double* t2d = nullptr;  // [YT][XT][TILE2]
  const double* t1d = nullptr;  // [XT][TILE1]
};

// Timing of the LU's kernels by HIP events (lu.hip): the whole factorization and solve, the
// trailing-update GEMMs with their algorithmic flops (8 m n k each), and the panels that needed
// interchanges (the pivoting fallback).
struct LuStats {
  double total_ms = 0.0, gemm_ms = 0.0, gemm_flops = 0.0;
  double refine_ms = 0.0;  // mixed precision: residuals, corrections and the FP64 store
  double first_eta = 0.0;  // mixed precision: the largest backward error of a first solution
  double big_ms = 0.0, big_flops = 0.0;  // the launches of the 128 x 128 split-FP16 GEMM tiles
  long big_launches = 0;
  long gemm_launches = 0;
  long fallback_panels = 0;
};

// ------------------------------------------------------------------ host helpers
struct Error {
  int status;
  std::string msg;
};

void set_error(int status, const std::string& msg, const char* where);

#define HBG_HIP(call)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) throw ::hbg::Error{6, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

// Host-side construction of the tables (green_tables.hip): node values on the device (or the host
// without a device), fits and monomial conversion on the host.
void build_wave_tables(bool on_device, std::vector<double>& t2d, std::vector<double>& t1d);
// The node evaluators, exposed for the CPU tests of the table builder.
void v_vx_node(double X, double Y, double* out2);
void line_node(double X, double* out4);

}  // namespace hbg
