// The C ABI of include/hydrobem_amd.h: sessions, geometry, the frequency loop on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hydrobem_amd.h"
#include "green_fin_dev.h"
#include "hbg_internal.h"
#include "surrogate.h"

namespace hbg {

// kernels (statics.hip, frequency.hip, lu.hip)
void launch_statics(const PanelD* P, const PanelD* Pf, int N, int Nsrc, int logint, double tol, double depth,
                    const double* gl, StaticD* st, int2* queue, unsigned* qcount, unsigned qcap, double2* inf,
                    hipStream_t stream);
void launch_forces_sol(const PanelD* P, const int* body, const double* nj, int N, int nrad, int H, double w,
                       double rho, const double2* sol, int lds, const double2* phiI, const double2* dphiI, double* out,
                       hipStream_t stream);
void launch_symmetry_combine(const double2* Sg, const double2* Ag, int N, int G, const double* chi, int nchar, double2* S,
                             double2* M, size_t mstride, int ldm, float2* A32, hipStream_t stream);
void launch_symmetry_project(const double2* u, int N, int G, int R, const double* chi, double2* uc,
                             hipStream_t stream);
void launch_symmetry_recombine(const double2* sol, size_t mstride, int lds, int N, int G, int R, const double* chi,
                               double2* phi, hipStream_t stream);
void launch_limit_matrices(const StaticD* st, const double2* inf, int N, int Nsrc, double2* S, double2* A,
                           hipStream_t stream);
void launch_assemble_fin(const PanelD* P, const PanelD* Psrc, int N, const FinFreq* f, FinTables F, WaveTables T,
                         double logc, const StaticD* st, double2* S, double2* A, float2* A32, int lda,
                         double free_term, hipStream_t stream, int images = 1, double wg = 0.0);
void launch_fin_points_eval(const FinFreq* f, FinTables F, WaveTables T, const double* R, const double* z,
                            const double* zeta, int n, double* out, hipStream_t stream);
size_t fin_tables_work();
void launch_fin_tables(const FinFreq& f, WaveTables T, const double* g16, double* phi1, double* phi2,
                       const double* M1dev, const PanelD* P, int N, double4* pf, double2* pcs, double* work,
                       hipStream_t stream);
FinFreq make_fin_freq(double h, double K);
double dispersion_root(double H, int n);
std::vector<double> fit_matrix(int M);
std::vector<double> build_k_table();
void launch_kochin(const PanelD* P, const int* body, const double* nj, int N, int nsol, int nrad, int H, double k,
                   double depth, int NK, const double2* sol, const double2* phiI, double2* out, hipStream_t stream);
void launch_field_values(const PanelD* Pf, const PanelD* P, const int* body, const double* nj, int Np, int N, int nsol,
                         int nrad, int H, double w, double K, double kw, double depth, const FinFreq* f, FinTables F,
                         double logc, WaveTables T, double grav, const double* heads, int walls, const StaticD* stf,
                         double2* Sx, double2* Ax, const double2* sol, const double2* phiI, double2* out,
                         hipStream_t stream);
void launch_adaptive(const PanelD* P, const PanelD* Pf, int N, const double* gl, StaticD* st, const int2* queue,
                     unsigned count, hipStream_t stream);
void launch_assemble(const PanelD* P, const PanelD* Psrc, int N, double K, double logc, WaveTables T,
                     const StaticD* st, double2* S, double2* A, float2* A32, int lda, double free_term,
                     hipStream_t stream, int images = 1, double wg = 0.0);
void launch_wave_points(WaveTables T, const double* X, const double* Y, long n, double* out, hipStream_t stream);
void launch_boundary(const PanelD* P, const int* body, const double* nj, int N, int nrad, int H, double w, double k,
                     double grav, double depth, const double* heads, double2* phiI, double2* dphiI, double2* u,
                     hipStream_t stream, int walls = 0);
void launch_rhs_batched(const double2* S, size_t sstride, int N, int lds, const double2* u, size_t ustride, int R,
                        double2* M, size_t mstride, int ldm, double2* part, int nsys, hipStream_t stream);
void launch_rhs(const double2* S, int N, int lds, const double2* u, int R, double2* M, int ldm, double2* part,
                hipStream_t stream);
int rhs_chunks(int N);
void launch_pad_system(double2* M, float2* M32, int n, int nz, int R, int rz, hipStream_t stream);
void launch_forces(const PanelD* P, const int* body, const double* nj, int N, int nrad, int H, double w, double rho,
                   const double2* M, int ldm, const double2* phiI, const double2* dphiI, double* out,
                   hipStream_t stream);
int lu_solve_device(double2* M, int n, int lda, int nrhs, hipStream_t stream, LuStats* stats);
void lu_solve_batched(double2* M, size_t mstride, int batch, int n, int lda, int nrhs, hipStream_t stream,
                      LuStats* stats, int* status, const int* nsys = nullptr);
int lu_solve_pivoted(double2* M, int n, int lda, int nrhs, hipStream_t stream, LuStats* stats);
size_t lu_region_workspace_bytes(int n);
int lu_solve_region(double2* M, int n, int lda, int nrhs, void* region, size_t region_bytes, int pivoted,
                    hipStream_t stream);
size_t noalloc_scratch_bytes();
void* noalloc_scratch_setup(int device);
// the FP32 far-field assembly (assembly32.hip)
void build_tables32(const WaveTables& T64, const double* ktab64, float* t2d32, float* t1d32, float* ktab32,
                    hipStream_t stream);
size_t tables32_t2d_elems();
size_t tables32_t1d_elems();
size_t tables32_ktab_elems();
size_t fin32_phi1_elems();
size_t fin32_phi2_elems();
size_t panel32_bytes();
size_t finfreq32_bytes();
double near_ratio();
std::vector<double> fin32_truncation();
void launch_fin_tables32(const double* phi1, const double* phi2, const double* dP, float* phi1_32, float* phi2_32,
                         const double4* pf, const double2* pcs, int N, float4* pf32, float2* pcs32,
                         hipStream_t stream);
void launch_panels32(const PanelD* P, int N, void* out, hipStream_t stream);
void launch_statics_list(const PanelD* P, int N, const int2* ent, int nnz, int logint, double tol, double depth,
                         const double* gl, StaticD* st, int2* queue, unsigned* qcount, unsigned qcap,
                         hipStream_t stream);
size_t tier_rules_elems(int N);
void launch_statics_far32(const PanelD* P, int N, int logint, double tol, double depth, const double* gl,
                          float4* rules, float4* st32, hipStream_t stream);
void launch_assemble32(const void* P32, int N, double K, double logc, const float* t2d32, const float* t1d32,
                       const float4* st32, float2* S, float2* A, float2* A2, int lda, double free_term,
                       hipStream_t stream);
void launch_assemble_fin32(const void* P32, int N, const void* f32, const float* phi1_32, const float* phi2_32,
                           const float* ktab32, const float4* pf32, const float2* pcs32, const float* t2d32,
                           const float* t1d32, double logc, const float4* st32, float2* S, float2* A, float2* A2,
                           int lda, double free_term, hipStream_t stream);
void launch_near64(const PanelD* P, int N, const int2* ent, int nnz, bool finite, double K, const FinFreq* f,
                   FinTables F, WaveTables T, double logc, const StaticD* st, float2* S, float2* A, float2* A2,
                   int lda, double free_term, double2* dS, double2* dA, hipStream_t stream);
void launch_rhs32(const float2* S, int N, const double2* u, int R, double2* B, double2* part, const int* rowptr,
                  const int2* ent, const double2* dS, hipStream_t stream);
void launch_wave_points32(const float* t2d32, const float* t1d32, const double* X, const double* Y, int n,
                          double* out, hipStream_t stream);
void launch_fin_points32(const void* f32, const float* phi1_32, const float* phi2_32, const float* ktab32,
                         const float4* pf32, const float2* pcs32, const float* t2d32, const float* t1d32,
                         const double* R, const double* z, const double* zeta, int n, double* out,
                         hipStream_t stream);
void* make_finfreq32(const FinFreq& f, void* host_out);
void mixed32_solve_batched(float2* M32, const float2* A32r, double2* B, const int* rowptr, const int2* ent,
                           const double2* dA, size_t dstride, double2* X, float2* D, double* part, double2* part2,
                           double* rscale, int batch, int n, int R, double tol, int max_iter, hipStream_t stream,
                           LuStats* stats, int* status, int* iters, double* eta, int* active);
void mixed_solve_batched(double2* M64, float2* M32, double2* X, float2* D, double* part, double2* part2,
                         double* rscale, int batch, int n, int R, bool a32_written, double tol, int max_iter,
                         hipStream_t stream, LuStats* stats, int* status, int* iters, double* eta,
                         const int* nsys, int* active);
int residual_blocks(int n);
size_t residual_max_elems(int n);
size_t residual_part_elems(int n, int R);
void launch_limit_fin_matrices(const PanelD* P, int N, int Nsrc, double h, bool zero, const double* ktab,
                               const StaticD* st, const double2* inf, double2* S, double2* A, hipStream_t stream);
void launch_interior_identity(const PanelD* P, int N, int nchar, double2* S, double2* M, size_t mstride,
                              hipStream_t stream);
void launch_limit_points(double h, bool zero, const double* ktab, int n, const double* R, const double* z,
                         const double* zeta, double* out, hipStream_t stream);
void launch_assemble_fin_nn(const PanelD* P, int N, const FinFreq* f, FinTables F, WaveTables T, double logc,
                            const StaticD* st, double2* S, double2* A, float2* A32, int lda, const NnDevice& nn,
                            int2* queue, unsigned* qcount, double rmin, hipStream_t stream, int images);
void launch_remainder(const FinFreq* f, FinTables F, WaveTables T, int n, const double* R, const double* z,
                      const double* zeta, double* out, hipStream_t stream);
void nn_eval(const NnDevice& nn, const float4* X, int n, float4* Y, hipStream_t stream);
double nn_gradient(const float* P_host, const float* X_host, const float* Y_host, int B, float* G_host,
                   hipStream_t st);
void nn_train(unsigned long long seed, int n_train, int n_hold, int epochs, int batch, float lr0, const WaveTables& T,
              const double* ktab, const double* fitm, const double* g16, std::vector<float>& P, double* report,
              NnDevice& nn, hipStream_t st);
void launch_motions(int F, int D, int H, const double* omegas, const double* M, const double* C, const double* Bx,
                    const double* A, const double* B, const double2* X, double2* out, int* status,
                    hipStream_t stream);
int motions_max_dof();
int motions_max_headings();
int lu_autotune(int precision64, int n, int nrhs, int max_batch, size_t mem_bytes, hipStream_t stream, double* report,
                int max_rows, int* rows);
void lu_set_graphs(int on);
void lu_graphs_note_free();  // device memory was freed: the LU's captured graphs are not replayed
int lu_graphs();
void lu_set_sub2(int q);
void lu_set_sub2_32(int q);
int lu_sub2();
int lu_sub2_32();
void lu_set_sub(int q);
void lu_set_sub32(int q);
int lu_sub();
int lu_sub32();
void lu_set_gemm_variant(int v);
int lu_gemm_variant();
void lu_set_gemm_variant32(int v);
int lu_gemm_variant32();
void lu_set_depth32(int q);
int lu_depth32();
void lu_set_depth(int q);
int lu_depth();
int write_wamit(const char* stem, int overwrite, double rho, double g, double L, const double* omegas, int F,
                const double* heads, int H, int nb, const double* A, const double* B, const double* HK,
                const double* X, const double* FK, const double* SC, const double* RAO, const char* stamp,
                std::string& err, int P = 0, const double* pts = nullptr,
                const double* frad = nullptr, const double* fdif = nullptr);

namespace {

thread_local std::string g_msg, g_loc;

struct V3 {
  double x[3];
  double& operator[](int i) { return x[i]; }
  double operator[](int i) const { return x[i]; }
};
V3 v3(double a, double b, double c) { return V3{{a, b, c}}; }
V3 operator+(V3 a, V3 b) { return v3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
V3 operator-(V3 a, V3 b) { return v3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
V3 operator*(double s, V3 a) { return v3(s * a[0], s * a[1], s * a[2]); }
double dot(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 cross(V3 a, V3 b) { return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]); }
double norm(V3 a) { return std::sqrt(dot(a, a)); }

// The panel rule of geometry.rs:120-199: projection onto the plane through the side midpoints
// (normal d1 x d2 of the diagonals), distinct vertices (coincidence 1e-6 of the diameter),
// strict convexity (sine 1e-6), area and centroid by the fan of triangles. Returns "" or the reason.
std::string make_panel(const double* v12, V3* proj, V3* poly, int& m, V3& normal, V3& centroid, double& area) {
  for (int i = 0; i < 12; ++i)
    if (!std::isfinite(v12[i])) return "non-finite vertex coordinate";
  V3 raw[4];
  for (int i = 0; i < 4; ++i) raw[i] = v3(v12[3 * i], v12[3 * i + 1], v12[3 * i + 2]);
  double diameter = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) diameter = std::max(diameter, norm(raw[i] - raw[j]));
  if (diameter == 0.0) return "fewer than three distinct vertices";
  const double tol = 1e-6 * diameter;
  int keep[4];
  m = 0;
  for (int i = 0; i < 4; ++i)
    if (norm(raw[i] - raw[(i + 1) % 4]) > tol) keep[m++] = i;
  if (m < 3) return "fewer than three distinct vertices";
  for (int a = 0; a < m; ++a)
    for (int b = a + 1; b < m; ++b)
      if (norm(raw[keep[a]] - raw[keep[b]]) <= tol)
        return "projected polygon is self-intersecting (coincident non-adjacent vertices)";
  V3 c = 0.25 * (raw[0] + raw[1] + raw[2] + raw[3]);
  V3 n = cross(raw[2] - raw[0], raw[3] - raw[1]);
  const double nn = norm(n);
  if (0.5 * nn <= 1e-6 * diameter * diameter) return "zero area (degenerate or self-intersecting polygon)";
  V3 nh = (1.0 / nn) * n;
  for (int i = 0; i < 4; ++i) proj[i] = raw[i] - dot(raw[i] - c, nh) * nh;
  for (int a = 0; a < m; ++a) poly[a] = proj[keep[a]];
  for (int a = 0; a < m; ++a) {
    V3 e1 = poly[(a + 1) % m] - poly[a];
    V3 e2 = poly[(a + 2) % m] - poly[(a + 1) % m];
    const double s = dot(cross(e1, e2), nh) / (norm(e1) * norm(e2));
    if (!(s > 1e-6)) return "projected polygon is not strictly convex or is self-intersecting";
  }
  area = 0.0;
  V3 mom = v3(0, 0, 0);
  for (int a = 1; a < m - 1; ++a) {
    const double t = 0.5 * dot(cross(poly[a] - poly[0], poly[a + 1] - poly[0]), nh);
    area += t;
    mom = mom + (t / 3.0) * (poly[0] + poly[a] + poly[a + 1]);
  }
  centroid = (1.0 / area) * mom;
  normal = nh;
  return "";
}

// The panel record of the kernels (panel_integrals.rs:145-217 make_panel_geom: distinct vertices
// of the projected corners, radius, edge lengths and in-plane edge normals).
PanelD panel_record(const V3* q, V3 normal, V3 centroid, double area) {
  PanelD p;
  std::memset(&p, 0, sizeof(p));
  double diam = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) diam = std::max(diam, norm(q[i] - q[j]));
  V3 v[4];
  int nv = 0;
  for (int i = 0; i < 4; ++i)
    if (norm(q[i] - q[(i + 1) % 4]) > 1e-6 * diam) v[nv++] = q[i];
  double radius = 0.0;
  for (int i = 0; i < nv; ++i) radius = std::max(radius, norm(v[i] - centroid));
  for (int e = 0; e < nv; ++e) {
    V3 edge = v[(e + 1) % nv] - v[e];
    p.elen[e] = norm(edge);
    V3 em = (1.0 / p.elen[e]) * cross(edge, normal);
    for (int d = 0; d < 3; ++d) p.em[e][d] = em[d];
  }
  if (nv == 3) v[3] = v[2];
  double top = v[0][2];
  for (int i = 0; i < 4; ++i) {
    for (int d = 0; d < 3; ++d) {
      p.v[i][d] = v[i][d];
      p.q4[i][d] = q[i][d];
    }
    if (i < nv) top = std::max(top, v[i][2]);
  }
  for (int d = 0; d < 3; ++d) {
    p.c[d] = centroid[d];
    p.n[d] = normal[d];
  }
  p.area = area;
  p.radius = radius;
  p.nv = nv;
  p.top = top;
  return p;
}

// Gauss-Legendre nodes and weights of orders 1..16, [order][x 16 | w 16] (quadrature.rs:17-59)
std::vector<double> gauss_table() {
  std::vector<double> t(17 * 32, 0.0);
  for (int n = 1; n <= 16; ++n) {
    double* x = &t[n * 32];
    double* w = x + 16;
    for (int i = 0; i < (n + 1) / 2; ++i) {
      double z = std::cos(PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
      for (int it = 0; it < 100; ++it) {
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
          double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
          p0 = p1;
          p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        double dz = p1 / dp;
        z -= dz;
        if (std::fabs(dz) <= 1e-16) break;
      }
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      x[i] = -z;
      x[n - 1 - i] = z;
      w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
  }
  return t;
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void ensure(size_t count) {
    if (count <= n && p) return;
    if (p) {
      HBG_HIP(hipFree(p));
      lu_graphs_note_free();
    }
    p = nullptr;
    HBG_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) {
      (void)hipFree(p);
      lu_graphs_note_free();
    }
    p = nullptr;
    n = 0;
  }
};

// process-wide wave-kernel tables of a device
struct TablesOnDevice {
  int device = -1;
  double* t2d = nullptr;
  double* t1d = nullptr;
  double* ktab = nullptr;  // K0, K1 tiles of the finite-depth series
  double* fitm = nullptr;  // node values -> monomial coefficients of the Phi tiles
  // the FP32 tables of the precision path, truncated from the FP64 ones (built on first use)
  float* t2d32 = nullptr;
  float* t1d32 = nullptr;
  float* ktab32 = nullptr;
  double* fin32P = nullptr;  // the Phi tiles' truncation matrix
};
TablesOnDevice& tables_for(int device) {
  static TablesOnDevice tabs[16];
  TablesOnDevice& t = tabs[device & 15];
  if (t.device == device) return t;
  std::vector<double> h2, h1;
  build_wave_tables(true, h2, h1);
  HBG_HIP(hipMalloc(&t.t2d, h2.size() * sizeof(double)));
  HBG_HIP(hipMalloc(&t.t1d, h1.size() * sizeof(double)));
  HBG_HIP(hipMemcpy(t.t2d, h2.data(), h2.size() * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(t.t1d, h1.data(), h1.size() * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> kt = build_k_table(), fm = fit_matrix(FM);
  HBG_HIP(hipMalloc(&t.ktab, kt.size() * sizeof(double)));
  HBG_HIP(hipMalloc(&t.fitm, fm.size() * sizeof(double)));
  HBG_HIP(hipMemcpy(t.ktab, kt.data(), kt.size() * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(t.fitm, fm.data(), fm.size() * sizeof(double), hipMemcpyHostToDevice));
  t.device = device;
  return t;
}

// The FP32 tables of the device (the tables of tables_for(device) truncated), once per device.
TablesOnDevice& tables32_for(int device, hipStream_t stream) {
  TablesOnDevice& t = tables_for(device);
  if (t.t2d32) return t;
  float *a = nullptr, *b = nullptr, *c = nullptr;
  double* d = nullptr;
  HBG_HIP(hipMalloc(&a, tables32_t2d_elems() * sizeof(float)));
  HBG_HIP(hipMalloc(&b, tables32_t1d_elems() * sizeof(float)));
  HBG_HIP(hipMalloc(&c, tables32_ktab_elems() * sizeof(float)));
  const std::vector<double> P = fin32_truncation();
  HBG_HIP(hipMalloc(&d, P.size() * sizeof(double)));
  HBG_HIP(hipMemcpy(d, P.data(), P.size() * sizeof(double), hipMemcpyHostToDevice));
  WaveTables T;
  T.t2d = t.t2d;
  T.t1d = t.t1d;
  build_tables32(T, t.ktab, a, b, c, stream);
  t.t2d32 = a;
  t.t1d32 = b;
  t.ktab32 = c;
  t.fin32P = d;
  return t;
}

}  // namespace

void set_error(int status, const std::string& msg, const char* where) {
  (void)status;
  g_msg = msg;
  g_loc = where ? where : "";
}

}  // namespace hbg

using namespace hbg;

struct HbgSession {
  int device = 0;
  hipStream_t stream = nullptr;     // assembly, boundary values, S U; everything outside a solve
  hipStream_t lu_stream = nullptr;  // a solve's LU, refinement and forces, beside the next batch's assembly
  hipEvent_t assembled[2] = {};     // the assembly of the batch in each half of the buffers is done
  double rho = 1000.0, grav = 9.80665, depth = INFINITY;
  int N = 0, nb = 0;  // N: the panels set (a symmetric body: its sector); the device arrays hold G N
  // symmetry planes (lo_geometry.rs:25-35): G = 1, 2 or 4 reflections in the order e, sigma_x
  // (x -> -x), sigma_y, sigma_x sigma_y (those present), chi[c][g] the characters
  int sym_x = 0, sym_y = 0, G = 1, Gs = 1;  // G reflections, Gs solved characters per frequency
  // walls (lo_geometry.rs:947-1031; bit 0: x = 0, bit 1: y = 0): the reflections are then those
  // of the walls, the image panels move as the mirror image of the body (their n_j copied), only
  // the character even under every wall is solved and the forces are those on the real panels
  int walls = 0;
  std::vector<std::array<double, 2>> refl{{1.0, 1.0}};
  std::vector<double> chi{1.0}, chisym{1.0};  // [Gs][G], [Gs][Gs]
  std::vector<double> raw_vertices, raw_centres;  // the panels as set, to rebuild the images
  std::vector<int> raw_body;
  std::vector<char> raw_interior;  // per panel: 1 interior free surface, 2 lid (after the body panels)
  std::vector<double> raw_eps;     // per panel: a lid panel's damping factor (empty: none)
  int NL = 0;                      // lid panels among the N
  int NI = 0;                      // panels of the interior free surface among the N (all images' sector)
  std::vector<PanelD> panels;
  std::vector<int> body;
  std::vector<double> nj;
  DevBuf<PanelD> d_panels;
  DevBuf<int> d_body;
  DevBuf<double> d_nj, d_gl, d_heads, d_out;
  DevBuf<StaticD> d_st;
  DevBuf<int2> d_queue;
  DevBuf<unsigned> d_qcount;
  DevBuf<double2> d_S, d_M, d_phiI, d_dphiI, d_u, d_part;
  // mixed precision: the complex64 systems, the FP64 solutions, the complex64 corrections, norms
  DevBuf<float2> d_M32, d_D32;
  DevBuf<double2> d_X64;
  DevBuf<double> d_rpart, d_rscale;
  DevBuf<double> d_chi, d_chisym;          // the solved characters [Gs][G], [Gs][Gs]
  DevBuf<double2> d_Sg, d_Ag, d_uc, d_phif;  // symmetric solves: an image block, projected u, phi
  DevBuf<int> d_active;  // systems still refined (mixed precision)
  DevBuf<double2> d_rpart2;
  NnDevice nn;          // the learned surrogate of the finite-depth evanescent remainder
  double nn_rmin = NN_RMIN;  // the network takes the pairs with nn_rmin <= R/h <= 20 (HBG_TUNE_NN_RMIN)
  DevBuf<int2> d_nnq;   // pairs left to the FP64 evaluator by the surrogate's assembly
  DevBuf<unsigned> d_nnc;
  int max_batch = 0;  // frequencies in flight (0: as many as fit, at most 48)
  std::vector<hipEvent_t> bev;
  bool statics_valid = false;
  bool stinf_valid = false;  // d_stinf holds the omega = infinity row of the statics
  DevBuf<double2> d_stinf;
  // field points (lo_frequency.rs:1392-1495): the points as records (their centroids), their
  // statics rows [N][np], the per-frequency influence S_x, A_x [N][np], the values of the batch,
  // and those of the last solve on the host ([F][np][nrad + 3 H] complex)
  int np = 0;
  std::vector<double> fpts;  // the points [np][3]
  DevBuf<PanelD> d_fpts;
  // finite depth: the panels and then the points (their factors of the tables at N + k), and
  // tables of their own (the field values of a batch are taken beside the next batch's assembly)
  DevBuf<PanelD> d_fcomb;
  DevBuf<double> d_phi1f, d_phi2f, d_fworkf;
  DevBuf<double4> d_pff;
  DevBuf<double2> d_pcsf;
  DevBuf<StaticD> d_stf;
  bool stf_valid = false;
  int stf_logint = -1;
  double stf_tol = -1.0;
  DevBuf<double2> d_Sx, d_Ax, d_fout;
  std::vector<double2> field_vals;
  int field_F = -1, field_H = 0, field_nrad = 0;
  // Kochin functions (lo_frequency.rs:1279-1372): NK angles, the batch's values [2][NK][R] per
  // frequency, those of the last solve on the host ([F][2][NK][R] complex)
  int nk = 0;
  DevBuf<double2> d_kout;
  std::vector<double2> kochin_vals;
  int kochin_F = -1, kochin_H = 0, kochin_nrad = 0;
  int st_logint = -1;
  double st_tol = -1.0;
  WaveTables T;
  // finite depth: per-frequency tiles and per-panel factors
  DevBuf<double> d_phi1, d_phi2, d_fwork;
  DevBuf<double4> d_pf;
  DevBuf<double2> d_pcs;
  DevBuf<FinFreq> d_ff;
  DevBuf<FinFreq> d_ffs;  // a solve's FinFreq of every ordinary frequency, uploaded once
  const double* ktab = nullptr;
  const double* fitm = nullptr;
  // profiling (hbg_set_profiling): GPU milliseconds of the stages, by HIP events on the stream
  bool profiling = false;
  double prof[HBG_PROFILE_COUNTERS] = {};
  hipEvent_t ev[8] = {};
  void* scratch = nullptr;  // this device's address of the allocation-free entry point's array
  // the FP32 far-field assembly (HBG_PRECISION_MIXED32): per geometry the FP32 panel records,
  // statics and the list of near entries (rows in order: rowptr [N + 1], (row, column) entries);
  // per frequency the FP32 tables and factors; the complex64 S, the FP64 right-hand sides and
  // solutions [system][N R], the near entries' FP64 differences (dS shared, dA per system)
  DevBuf<unsigned char> d_panels32;
  DevBuf<float4> d_st32, d_rules;
  bool st32_valid = false;  // d_st32 and the FP64 records of the near entries, for st32_logint, st32_tol
  int st32_logint = -1;
  double st32_tol = -1.0;
  bool near_valid = false;
  int near_nnz = 0;
  std::vector<int2> near_ent;  // host copy of the near entries (row, column)
  DevBuf<int> d_near_rowptr;
  DevBuf<int2> d_near_ent;
  DevBuf<float> d_phi1_32, d_phi2_32;
  DevBuf<float4> d_pf32;
  DevBuf<float2> d_pcs32;
  DevBuf<unsigned char> d_ffs32;
  DevBuf<float2> d_S32, d_A32r;  // S; the unfactored copy of A per system (the residuals)
  DevBuf<double2> d_B64, d_dS, d_dA;
};

namespace {

HbgStatus fail(int status, const std::string& msg, const char* where) {
  set_error(status, msg, where);
  return status;
}

#define GUARD_BEGIN try {
#define GUARD_END(where)                                             \
  }                                                                  \
  catch (const ::hbg::Error& e) {                                    \
    return fail(e.status, e.msg, where);                             \
  }                                                                  \
  catch (const std::exception& e) {                                  \
    return fail(HBG_ERR_INTERNAL, std::string("internal: ") + e.what(), where); \
  }                                                                  \
  catch (...) {                                                      \
    return fail(HBG_ERR_INTERNAL, "internal error", where);          \
  }

void use_device(HbgSession* s) { HBG_HIP(hipSetDevice(s->device)); }

void ensure_statics(HbgSession* s, int logint, double tol, bool want_inf = false) {
  if (s->statics_valid && s->st_logint == logint && s->st_tol == tol && (!want_inf || s->stinf_valid)) return;
  const int N = s->N, NF = s->G * N;
  s->d_st.ensure((size_t)N * NF);
  if (want_inf) s->d_stinf.ensure((size_t)N * NF);
  unsigned qcap = (unsigned)std::min<size_t>((size_t)N * 64, (size_t)1 << 28);
  if (s->profiling) HBG_HIP(hipEventRecord(s->ev[0], s->stream));
  for (int attempt = 0; attempt < 2; ++attempt) {
    s->d_queue.ensure(qcap);
    s->d_qcount.ensure(1);
    HBG_HIP(hipMemsetAsync(s->d_qcount.p, 0, sizeof(unsigned), s->stream));
    launch_statics(s->d_panels.p, s->d_panels.p, N, NF, logint, tol, std::isfinite(s->depth) ? s->depth : 0.0, s->d_gl.p,
                   s->d_st.p, s->d_queue.p, s->d_qcount.p, qcap, want_inf ? s->d_stinf.p : nullptr, s->stream);
    unsigned count = 0;
    HBG_HIP(hipMemcpyAsync(&count, s->d_qcount.p, sizeof(unsigned), hipMemcpyDeviceToHost, s->stream));
    HBG_HIP(hipStreamSynchronize(s->stream));
    if (count <= qcap) {
      launch_adaptive(s->d_panels.p, s->d_panels.p, N, s->d_gl.p, s->d_st.p, s->d_queue.p, count, s->stream);
      if (s->profiling) HBG_HIP(hipEventRecord(s->ev[1], s->stream));
      HBG_HIP(hipStreamSynchronize(s->stream));
      if (s->profiling) {
        float ms = 0.0f;
        HBG_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
        s->prof[0] += ms;
        s->prof[11] += 1.0;
      }
      s->statics_valid = true;
      s->stinf_valid = want_inf;
      s->st_logint = logint;
      s->st_tol = tol;
      return;
    }
    qcap = count;
  }
  throw Error{HBG_ERR_INTERNAL, "statics: adaptive queue overflow"};
}

// The statics rows of the field points (the same kernels, the points as field records).
void ensure_field_statics(HbgSession* s, int logint, double tol) {
  if (s->stf_valid && s->stf_logint == logint && s->stf_tol == tol) return;
  const int N = s->G * s->N, np = s->np;  // sources: the panels and their images
  s->d_stf.ensure((size_t)N * np);
  unsigned qcap = (unsigned)std::min<size_t>((size_t)np * 64 + 64, (size_t)1 << 28);
  for (int attempt = 0; attempt < 2; ++attempt) {
    s->d_queue.ensure(qcap);
    s->d_qcount.ensure(1);
    HBG_HIP(hipMemsetAsync(s->d_qcount.p, 0, sizeof(unsigned), s->stream));
    launch_statics(s->d_panels.p, s->d_fpts.p, np, N, logint, tol, std::isfinite(s->depth) ? s->depth : 0.0, s->d_gl.p,
                   s->d_stf.p, s->d_queue.p, s->d_qcount.p, qcap, nullptr, s->stream);
    unsigned count = 0;
    HBG_HIP(hipMemcpyAsync(&count, s->d_qcount.p, sizeof(unsigned), hipMemcpyDeviceToHost, s->stream));
    HBG_HIP(hipStreamSynchronize(s->stream));
    if (count <= qcap) {
      launch_adaptive(s->d_panels.p, s->d_fpts.p, np, s->d_gl.p, s->d_stf.p, s->d_queue.p, count, s->stream);
      if (std::isfinite(s->depth)) {
        s->d_fcomb.ensure((size_t)N + np);
        HBG_HIP(hipMemcpyAsync(s->d_fcomb.p, s->d_panels.p, (size_t)N * sizeof(PanelD), hipMemcpyDeviceToDevice,
                               s->stream));
        HBG_HIP(hipMemcpyAsync(s->d_fcomb.p + N, s->d_fpts.p, (size_t)np * sizeof(PanelD), hipMemcpyDeviceToDevice,
                               s->stream));
      }
      HBG_HIP(hipStreamSynchronize(s->stream));
      s->stf_valid = true;
      s->stf_logint = logint;
      s->stf_tol = tol;
      return;
    }
    qcap = count;
  }
  throw Error{HBG_ERR_INTERNAL, "the adaptive queue of the field points did not fit"};
}

// the depths of the panels must lie in the water: -h <= z < 0 at the centroids
void check_depth(HbgSession* s) {
  if (!std::isfinite(s->depth)) return;
  for (int j = 0; j < s->N; ++j)
    if (s->panels[j].c[2] < -s->depth)
      throw Error{HBG_ERR_VALUE, "panel " + std::to_string(j + 1) + ": centroid below the sea bed"};
}

FinTables fin_tables(HbgSession* s) {
  FinTables F;
  F.phi1 = s->d_phi1.p;
  F.phi2 = s->d_phi2.p;
  F.ktab = s->ktab;
  F.pf = s->d_pf.p;
  F.pcs = s->d_pcs.p;
  return F;
}

// S and A of the frequency K into S (ld N) and M (ld N): infinite depth, or finite depth with the
// per-frequency tables built first; returns the wave number of the incident wave
// S and A (and A32) with leading dimension lda (at least N; S in s->d_S, lda x N)
// The matrices of one frequency: S (d_S) and A into Mout (leading dimension lda), A32 the
// complex64 copy (nullable). A symmetric body (G > 1): the G character systems at Mout + c mstride
// (leading dimension lda) and S_c at d_S + c N N, from its image blocks (scratch d_Sg, d_Ag).
// dff: the frequency's FinFreq already on the device (finite depth; nullptr: uploaded here, which
// waits for the stream, the host copy being pageable).
double assemble_frequency(HbgSession* s, double K, double logc, double2* Mout, bool surrogate = false,
                          float2* A32 = nullptr, int lda = 0, size_t mstride = 0, const FinFreq* dff = nullptr) {
  const int N = s->N, G = s->G;
  if (lda <= 0) lda = N;
  if (G > 1 && !mstride) throw Error{HBG_ERR_INTERNAL, "symmetry planes without the stride of the character systems"};
  if (G > 1) {
    s->d_Sg.ensure((size_t)N * N * G);
    s->d_Ag.ensure((size_t)N * N * G);
  }
  auto blocks = [&](auto one) {
    if (G == 1) {
      one(s->d_S.p, Mout, A32, 1, lda);
      return;
    }
    // the G image blocks (leading dimension N) in one launch, each into its own buffer, then the
    // G characters in one pass
    one(s->d_Sg.p, s->d_Ag.p, (float2*)nullptr, G, N);
    launch_symmetry_combine(s->d_Sg.p, s->d_Ag.p, N, G, s->d_chi.p, s->Gs, s->d_S.p, Mout, mstride, lda, A32,
                            s->stream);
  };
  if (!std::isfinite(s->depth)) {
    blocks([&](double2* S, double2* Ab, float2* a32, int images, int ld) {
      launch_assemble(s->d_panels.p, s->d_panels.p, N, K, logc, s->T, s->d_st.p, S, Ab, a32, ld, 2.0 * PI, s->stream,
                      images, std::sqrt(K / s->grav));
    });
    return K;
  }
  const FinFreq f = make_fin_freq(s->depth, K);
  s->d_phi1.ensure((size_t)FTS * FTW * FTILE);
  s->d_phi2.ensure((size_t)FTS * FTV * FTILE);
  s->d_pf.ensure(N);
  s->d_pcs.ensure((size_t)N * FIN_ROOTS);
  s->d_fwork.ensure(fin_tables_work());
  launch_fin_tables(f, s->T, s->d_gl.p + 16 * 32, s->d_phi1.p, s->d_phi2.p, s->fitm, s->d_panels.p, N, s->d_pf.p,
                    s->d_pcs.p, s->d_fwork.p, s->stream);
  if (!dff) {
    s->d_ff.ensure(1);
    HBG_HIP(hipMemcpyAsync(s->d_ff.p, &f, sizeof(FinFreq), hipMemcpyHostToDevice, s->stream));
    dff = s->d_ff.p;
  }
  const double nuh = K * s->depth;
  if (surrogate && s->nn.ready && nuh >= std::exp(NN_LNNU_LO) && nuh <= std::exp(NN_LNNU_HI)) {
    s->d_nnq.ensure((size_t)N * (N + 1) / 2 * G);  // every pair of every image block may fall back
    s->d_nnc.ensure(1);
    blocks([&](double2* S, double2* Ab, float2* a32, int images, int ld) {
      launch_assemble_fin_nn(s->d_panels.p, N, dff, fin_tables(s), s->T, logc, s->d_st.p, S, Ab, a32, ld, s->nn,
                             s->d_nnq.p, s->d_nnc.p, s->nn_rmin, s->stream, images);
    });
  } else
    blocks([&](double2* S, double2* Ab, float2* a32, int images, int ld) {
      launch_assemble_fin(s->d_panels.p, s->d_panels.p, N, dff, fin_tables(s), s->T, logc, s->d_st.p, S, Ab, a32, ld,
                          2.0 * PI, s->stream, images, std::sqrt(K / s->grav));
    });
  return f.k0;
}

// a group solve's batch takes the next system while its order is at least this share of the
// batch's first (largest) order
constexpr double GROUP_ORDER_RATIO = 0.8;
constexpr double MIXED_TOL = 1e-12;  // normwise backward error that ends the refinement
constexpr int MIXED_MAX_ITER = 10;
// the refinement of the FP32 assembly ends at 10^-g_mixed32_tol_exp (HBG_TUNE_MIXED32_TOL)
int g_mixed32_tol_exp = 10;

// ---- the FP32 far-field assembly (HBG_PRECISION_MIXED32)

void ensure_near_list(HbgSession* s);

// The FP32 panel records and statics of the geometry (G = 1) and its near entries (the pairs
// whose centroids are closer than near_ratio() (r_k + r_n)): FP32 records of every pair, by the
// Gauss tiers beyond two panel radii (statics_far32_kernel; 0 nearer), and the FP64 records of
// the near entries (all pairs' FP64 records when the FP64 statics are current already), which
// the near entries of the assembly take instead. Counted as the statics in the stage timing.
void ensure_fp32_geometry(HbgSession* s, int logint, double tol) {
  const int N = s->N;
  ensure_near_list(s);
  if (s->st32_valid && s->st32_logint == logint && s->st32_tol == tol) return;
  if (s->profiling) HBG_HIP(hipEventRecord(s->ev[0], s->stream));
  s->d_panels32.ensure((size_t)N * panel32_bytes());
  launch_panels32(s->d_panels.p, N, s->d_panels32.p, s->stream);
  const double depth = std::isfinite(s->depth) ? s->depth : 0.0;
  if (!(s->statics_valid && s->st_logint == logint && s->st_tol == tol)) {
    s->d_st.ensure((size_t)N * N);
    const unsigned qcap = (unsigned)std::max(s->near_nnz, 1);
    s->d_queue.ensure(qcap);
    s->d_qcount.ensure(1);
    HBG_HIP(hipMemsetAsync(s->d_qcount.p, 0, sizeof(unsigned), s->stream));
    launch_statics_list(s->d_panels.p, N, s->d_near_ent.p, s->near_nnz, logint, tol, depth, s->d_gl.p, s->d_st.p,
                        s->d_queue.p, s->d_qcount.p, qcap, s->stream);
    unsigned count = 0;
    HBG_HIP(hipMemcpyAsync(&count, s->d_qcount.p, sizeof(unsigned), hipMemcpyDeviceToHost, s->stream));
    HBG_HIP(hipStreamSynchronize(s->stream));
    if (count > qcap) throw Error{HBG_ERR_INTERNAL, "statics: adaptive queue overflow at the near entries"};
    launch_adaptive(s->d_panels.p, s->d_panels.p, N, s->d_gl.p, s->d_st.p, s->d_queue.p, count, s->stream);
  }
  s->d_rules.ensure(tier_rules_elems(N));
  s->d_st32.ensure((size_t)N * N);
  launch_statics_far32(s->d_panels.p, N, logint, tol, depth, s->d_gl.p, s->d_rules.p, s->d_st32.p, s->stream);
  if (s->profiling) HBG_HIP(hipEventRecord(s->ev[1], s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  if (s->profiling) {
    float ms = 0.0f;
    HBG_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
    s->prof[0] += ms;
    s->prof[11] += 1.0;
  }
  s->st32_valid = true;
  s->st32_logint = logint;
  s->st32_tol = tol;
}

// The near entries of the geometry, found on a uniform grid of the centroids, every row k's
// columns in increasing order, the diagonal included.
void ensure_near_list(HbgSession* s) {
  const int N = s->N;
  if (s->near_valid) return;
  const double ratio = near_ratio();
  double rmax = 0.0, lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int j = 0; j < N; ++j) {
    rmax = std::max(rmax, s->panels[j].radius);
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], s->panels[j].c[d]);
      hi[d] = std::max(hi[d], s->panels[j].c[d]);
    }
  }
  const double cell = std::max(2.0 * ratio * rmax, 1e-12);
  int dims[3];
  for (int d = 0; d < 3; ++d) dims[d] = std::max(1, std::min(1024, (int)((hi[d] - lo[d]) / cell) + 1));
  auto cell_of = [&](int j, int d) {
    return std::min(dims[d] - 1, std::max(0, (int)((s->panels[j].c[d] - lo[d]) / cell)));
  };
  std::vector<std::vector<int>> bins((size_t)dims[0] * dims[1] * dims[2]);
  for (int j = 0; j < N; ++j)
    bins[((size_t)cell_of(j, 2) * dims[1] + cell_of(j, 1)) * dims[0] + cell_of(j, 0)].push_back(j);
  std::vector<int> rowptr(N + 1, 0);
  std::vector<int2> ent;
  for (int k = 0; k < N; ++k) {
    const PanelD& pk = s->panels[k];
    std::vector<int> cols;
    const int ck[3] = {cell_of(k, 0), cell_of(k, 1), cell_of(k, 2)};
    for (int z = std::max(0, ck[2] - 1); z <= std::min(dims[2] - 1, ck[2] + 1); ++z)
      for (int y = std::max(0, ck[1] - 1); y <= std::min(dims[1] - 1, ck[1] + 1); ++y)
        for (int x = std::max(0, ck[0] - 1); x <= std::min(dims[0] - 1, ck[0] + 1); ++x)
          for (int n : bins[((size_t)z * dims[1] + y) * dims[0] + x]) {
            const PanelD& pn = s->panels[n];
            const double dx = pk.c[0] - pn.c[0], dy = pk.c[1] - pn.c[1], dz = pk.c[2] - pn.c[2];
            if (std::sqrt(dx * dx + dy * dy + dz * dz) < ratio * (pk.radius + pn.radius)) cols.push_back(n);
          }
    std::sort(cols.begin(), cols.end());
    for (int n : cols) ent.push_back(make_int2(k, n));
    rowptr[k + 1] = (int)ent.size();
  }
  s->near_nnz = (int)ent.size();
  s->near_ent = ent;
  s->d_near_rowptr.ensure(rowptr.size());
  s->d_near_ent.ensure(std::max<size_t>(ent.size(), 1));
  HBG_HIP(hipMemcpy(s->d_near_rowptr.p, rowptr.data(), rowptr.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!ent.empty())
    HBG_HIP(hipMemcpy(s->d_near_ent.p, ent.data(), ent.size() * sizeof(int2), hipMemcpyHostToDevice));
  s->near_valid = true;
}

// Whether every pair of a frequency K lies inside the FP32 tables (the FP32 kernel has no
// asymptotic forms): infinite depth X = K R < 48 and Y = K |z + zeta| < 48 over the body's extent;
// finite depth the same for the Newman form (R < h / 2) and k0 R < 48 for the series.
bool fp32_in_range(HbgSession* s, double K) {
  const int N = s->N;
  double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY}, zmin = 0.0;
  for (int j = 0; j < N; ++j) {
    for (int d = 0; d < 2; ++d) {
      lo[d] = std::min(lo[d], s->panels[j].c[d]);
      hi[d] = std::max(hi[d], s->panels[j].c[d]);
    }
    zmin = std::min(zmin, s->panels[j].c[2]);
  }
  const double Rmax = std::hypot(hi[0] - lo[0], hi[1] - lo[1]);
  const double lim = 0.999 * XT;
  if (!std::isfinite(s->depth)) return K * Rmax < lim && K * 2.0 * -zmin < lim;
  const FinFreq f = make_fin_freq(s->depth, K);
  return K * std::min(Rmax, 0.5 * s->depth) < lim && K * 2.0 * -zmin < lim && f.k0 * Rmax < lim;
}

// S (complex64, s->d_S32) and A (complex64, ld N) of the frequency K by the FP32 kernels, the near
// entries in FP64 (their differences into s->d_dS and dA); fi: the frequency's index in the
// solve's preloaded FinFreq arrays (finite depth). Returns the incident wave's wave number.
// A32r: the copy of A (ld N) that stays unfactored for the refinement's residuals.
double assemble_frequency32(HbgSession* s, double K, double logc, int fi, float2* A32, float2* A32r, double2* dA) {
  const int N = s->N;
  TablesOnDevice& t = tables32_for(s->device, s->stream);
  if (!std::isfinite(s->depth)) {
    launch_assemble32(s->d_panels32.p, N, K, logc, t.t2d32, t.t1d32, s->d_st32.p, s->d_S32.p, A32, nullptr, N,
                      2.0 * PI, s->stream);
    launch_near64(s->d_panels.p, N, s->d_near_ent.p, s->near_nnz, false, K, nullptr, FinTables{}, s->T, logc,
                  s->d_st.p, s->d_S32.p, A32, nullptr, N, 2.0 * PI, s->d_dS.p, dA, s->stream);
    // the refinement's copy of A (the LU factors A in place), in one coalesced pass
    HBG_HIP(hipMemcpyAsync(A32r, A32, (size_t)N * N * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
    return K;
  }
  const FinFreq f = make_fin_freq(s->depth, K);
  s->d_phi1.ensure((size_t)FTS * FTW * FTILE);
  s->d_phi2.ensure((size_t)FTS * FTV * FTILE);
  s->d_pf.ensure(N);
  s->d_pcs.ensure((size_t)N * FIN_ROOTS);
  s->d_fwork.ensure(fin_tables_work());
  launch_fin_tables(f, s->T, s->d_gl.p + 16 * 32, s->d_phi1.p, s->d_phi2.p, s->fitm, s->d_panels.p, N, s->d_pf.p,
                    s->d_pcs.p, s->d_fwork.p, s->stream);
  s->d_phi1_32.ensure(fin32_phi1_elems());
  s->d_phi2_32.ensure(fin32_phi2_elems());
  s->d_pf32.ensure(N);
  s->d_pcs32.ensure((size_t)N * FIN_ROOTS);
  launch_fin_tables32(s->d_phi1.p, s->d_phi2.p, t.fin32P, s->d_phi1_32.p, s->d_phi2_32.p, s->d_pf.p, s->d_pcs.p, N,
                      s->d_pf32.p, s->d_pcs32.p, s->stream);
  launch_assemble_fin32(s->d_panels32.p, N, s->d_ffs32.p + (size_t)fi * finfreq32_bytes(), s->d_phi1_32.p,
                        s->d_phi2_32.p, t.ktab32, s->d_pf32.p, s->d_pcs32.p, t.t2d32, t.t1d32, logc, s->d_st32.p,
                        s->d_S32.p, A32, nullptr, N, 2.0 * PI, s->stream);
  launch_near64(s->d_panels.p, N, s->d_near_ent.p, s->near_nnz, true, K, s->d_ffs.p + fi, fin_tables(s), s->T, logc,
                s->d_st.p, s->d_S32.p, A32, nullptr, N, 2.0 * PI, s->d_dS.p, dA, s->stream);
  // the refinement's copy of A (the LU factors A in place), in one coalesced pass
  HBG_HIP(hipMemcpyAsync(A32r, A32, (size_t)N * N * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
  return f.k0;
}

// Device bytes per frequency in flight: its N x (N + R) system (and its complex64 copy), the
// LU's per-panel inverses, the refinement's vectors.
size_t bytes_per_frequency(int N, int R, bool mixed) {
  size_t per = (size_t)N * (N + R) * sizeof(double2) + (size_t)N * 64 * sizeof(double2) * 2;
  if (mixed) per += (size_t)N * (N + R) * sizeof(float2) + (size_t)N * R * (sizeof(double2) + sizeof(float2));
  return per;
}

// Frequencies in flight: the session's setting, else as many as half the free device memory
// holds (copies of the batch buffers each), at most 48, at most the number of frequencies.
int batch_size(HbgSession* s, int N, int R, int F, bool mixed, int copies = 1) {
  const size_t per = bytes_per_frequency(N, R, mixed) * copies;
  int B = s->max_batch;
  if (B <= 0) {
    size_t free_b = 0, total_b = 0;
    HBG_HIP(hipMemGetInfo(&free_b, &total_b));
    B = (int)std::min<size_t>(48, std::max<size_t>(1, free_b / 2 / per));
  }
  return std::max(1, std::min(B, std::max(F, 1)));
}

void check_options(const HbgOptions* o) {
  if (o && o->precision != HBG_PRECISION_FP64 && o->precision != HBG_PRECISION_MIXED &&
      o->precision != HBG_PRECISION_MIXED32)
    throw Error{HBG_ERR_UNSUPPORTED, "precision not supported"};
  if (o && !(o->log_tolerance >= 0.0)) throw Error{HBG_ERR_VALUE, "log_tolerance must be >= 0"};
  if (o && o->surrogate != 0 && o->surrogate != 1) throw Error{HBG_ERR_VALUE, "surrogate must be 0 or 1"};
}

HbgOptions defaults(const HbgOptions* o) {
  HbgOptions d;
  d.precision = HBG_PRECISION_FP64;
  d.logarithmic = 1;
  d.log_tolerance = 1e-4;
  d.surrogate = 0;
  d.reserved = 0;
  return o ? *o : d;
}

}  // namespace

extern "C" {

const char* hbg_last_error_message(void) { return g_msg.c_str(); }
const char* hbg_last_error_location(void) { return g_loc.c_str(); }
const char* hbg_version(void) {
  return "hydrobem_amd 0.1.0 (gfx950; low-order frequency-domain solve, infinite and finite depth, FP64 and mixed "
         "precision)";
}

HbgStatus hbg_session_create(int32_t device, HbgSession** out) {
  if (!out) return fail(HBG_ERR_ARGUMENT, "out is null", "hbg_session_create");
  GUARD_BEGIN
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    throw Error{HBG_ERR_STATE, "no GPU device: the HIP library needs an MI355X (gfx950)"};
  if (device < 0 || device >= count) throw Error{HBG_ERR_ARGUMENT, "device index out of range"};
  std::unique_ptr<HbgSession> s(new HbgSession());
  s->device = device;
  HBG_HIP(hipSetDevice(device));
  // a blocking stream: ordered with the null stream, so the synchronous copies and fills of the
  // host code (hipMemcpy, hipMemset) complete before the session's kernels read their buffers
  HBG_HIP(hipStreamCreate(&s->stream));
  HBG_HIP(hipStreamCreate(&s->lu_stream));
  for (auto& e : s->assembled) HBG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (auto& e : s->ev) HBG_HIP(hipEventCreate(&e));
  s->scratch = noalloc_scratch_setup(device);
  TablesOnDevice& t = tables_for(device);
  s->T.t2d = t.t2d;
  s->T.t1d = t.t1d;
  s->ktab = t.ktab;
  s->fitm = t.fitm;
  std::vector<double> gl = gauss_table();
  s->d_gl.ensure(gl.size());
  HBG_HIP(hipMemcpy(s->d_gl.p, gl.data(), gl.size() * sizeof(double), hipMemcpyHostToDevice));
  *out = s.release();
  return HBG_OK;
  GUARD_END("hbg_session_create")
}

HbgStatus hbg_session_free(HbgSession* s) {
  if (!s) return HBG_OK;
  GUARD_BEGIN
  (void)hipSetDevice(s->device);
  (void)hipStreamSynchronize(s->stream);
  s->d_panels.release();
  s->d_body.release();
  s->d_nj.release();
  s->d_gl.release();
  s->d_heads.release();
  s->d_out.release();
  s->d_st.release();
  s->d_queue.release();
  s->d_qcount.release();
  s->d_S.release();
  s->d_M.release();
  s->d_phiI.release();
  s->d_dphiI.release();
  s->d_u.release();
  s->d_part.release();
  s->d_phi1.release();
  s->d_phi2.release();
  s->d_fwork.release();
  s->d_pf.release();
  s->d_pcs.release();
  s->d_ff.release();
  s->d_active.release();
  s->d_stinf.release();
  (void)hipStreamSynchronize(s->lu_stream);
  for (auto& e : s->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : s->assembled)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(s->lu_stream);
  (void)hipStreamDestroy(s->stream);
  delete s;
  return HBG_OK;
  GUARD_END("hbg_session_free")
}

HbgStatus hbg_set_environment(HbgSession* s, double density, double gravity, double depth) {
  if (!s) return fail(HBG_ERR_ARGUMENT, "session is null", "hbg_set_environment");
  if (!(density > 0.0) || !std::isfinite(density)) return fail(HBG_ERR_VALUE, "density must be > 0", "hbg_set_environment");
  if (!(gravity > 0.0) || !std::isfinite(gravity)) return fail(HBG_ERR_VALUE, "gravity must be > 0", "hbg_set_environment");
  if (!(depth > 0.0)) return fail(HBG_ERR_VALUE, "depth must be > 0", "hbg_set_environment");
  if (depth != s->depth) s->statics_valid = s->stf_valid = s->st32_valid = false;
  s->rho = density;
  s->grav = gravity;
  s->depth = depth;
  return HBG_OK;
}

// The panels as the caller set them, kept to rebuild the images when the planes or walls change.
struct RawPanels {
  std::vector<double> vertices, centres;
  std::vector<int> body;
  std::vector<char> interior;  // per panel: 1 interior free surface, 2 lid (after the body's), 0 body
  std::vector<double> eps;     // per panel: a lid panel's damping factor (empty: none)
  int N = 0, nb = 0, NI = 0;
};

// The reflection group of the symmetry planes and walls and its solved characters.
struct Group {
  int G = 1, Gs = 1;
  std::vector<std::array<double, 2>> refl{{1.0, 1.0}};
  std::vector<double> chi{1.0}, chisym{1.0};  // [Gs][G], [Gs][Gs]
};
static Group make_group(int sym_x, int sym_y, int walls);
static void set_geometry(HbgSession* s, const Group& grp, int sym_x, int sym_y, int walls, const RawPanels& raw);

static RawPanels raw_of(const HbgSession* s) {
  RawPanels r;
  r.vertices = s->raw_vertices;
  r.centres = s->raw_centres;
  r.body = s->raw_body;
  r.interior = s->raw_interior;
  r.eps = s->raw_eps;
  r.N = s->N;
  r.nb = s->nb;
  r.NI = s->NI;
  return r;
}

HbgStatus hbg_set_panels(HbgSession* s, const double* vertices, int32_t n_panels, const int32_t* body,
                         int32_t n_bodies, const double* rotation_centres) {
  if (!s || !vertices) return fail(HBG_ERR_ARGUMENT, "session or vertices is null", "hbg_set_panels");
  if (n_panels <= 0 || n_bodies <= 0) return fail(HBG_ERR_ARGUMENT, "n_panels and n_bodies must be > 0", "hbg_set_panels");
  GUARD_BEGIN
  RawPanels raw;
  raw.vertices.assign(vertices, vertices + 12 * (size_t)n_panels);
  raw.body.assign(n_panels, 0);
  for (int j = 0; j < n_panels; ++j) raw.body[j] = body ? body[j] : 0;
  raw.centres.assign(3 * (size_t)n_bodies, 0.0);
  if (rotation_centres) raw.centres.assign(rotation_centres, rotation_centres + 3 * (size_t)n_bodies);
  raw.interior.assign(n_panels, 0);
  raw.N = n_panels;
  raw.nb = n_bodies;
  // a refused panel leaves the session's geometry (and its statics) as they were
  set_geometry(s, make_group(s->sym_x, s->sym_y, s->walls), s->sym_x, s->sym_y, s->walls, raw);
  return HBG_OK;
  GUARD_END("hbg_set_panels")
}

// The device panels of a geometry: the panels as set and, with symmetry planes or walls, their
// images in the order of the group's reflections (geometry.rs:209-234: the vertices reflected,
// their order reversed where one reflection reverses the orientation), each with its body's
// modes. Everything is built and checked on the host first; the session (group, panels, raw
// arrays, device copies) changes only when every panel passed, and every result derived from the
// previous geometry (statics, field statics, field and Kochin values) is invalidated.
static void set_geometry(HbgSession* s, const Group& grp, int sym_x, int sym_y, int walls, const RawPanels& raw) {
  const int n_panels = raw.N, n_bodies = raw.nb, G = grp.G, NF = G * n_panels;
  std::vector<PanelD> panels(NF);
  std::vector<int> b(NF, 0);
  std::vector<double> nj(6 * (size_t)NF);
  for (int g = 0; g < G; ++g)
    for (int j = 0; j < n_panels; ++j) {
      const double sx = grp.refl[g][0], sy = grp.refl[g][1];
      if (g == 0 && walls)  // the body in the fluid: x >= 0 (a wall at x = 0), y >= 0 (y = 0)
        for (int q = 0; q < 4; ++q)
          for (int a = 0; a < 2; ++a)
            if ((walls >> a & 1) && raw.vertices[12 * (size_t)j + 3 * q + a] < 0.0)
              throw Error{HBG_ERR_VALUE, "panel " + std::to_string(j + 1) + ": behind the wall " +
                                             (a ? "y" : "x") + " = 0"};
      const bool inner = !raw.interior.empty() && raw.interior[j];
      double v[12];
      for (int q = 0; q < 4; ++q) {
        const int src = (sx * sy < 0.0 && q > 0) ? 4 - q : q;  // 0, 3, 2, 1 where the orientation flips
        v[3 * q] = sx * raw.vertices[12 * (size_t)j + 3 * src];
        v[3 * q + 1] = sy * raw.vertices[12 * (size_t)j + 3 * src + 1];
        v[3 * q + 2] = inner ? 0.0 : raw.vertices[12 * (size_t)j + 3 * src + 2];
      }
      V3 proj[4], poly[4], normal, centroid;
      int m;
      double area;
      std::string why = make_panel(v, proj, poly, m, normal, centroid, area);
      if (why.empty() && inner && normal[2] < 0.0) {
        // the interior free surface lies in z = 0 with its normal up (lo_geometry.rs:707-718):
        // a panel given clockwise from above is taken counter-clockwise
        for (int d = 0; d < 3; ++d) std::swap(v[3 + d], v[9 + d]);
        why = make_panel(v, proj, poly, m, normal, centroid, area);
      }
      if (!why.empty()) throw Error{HBG_ERR_VALUE, "panel " + std::to_string(j + 1) + ": " + why};
      if (!inner && !(centroid[2] < 0.0))
        throw Error{HBG_ERR_UNSUPPORTED,
                    "panel " + std::to_string(j + 1) + ": centroid not below z = 0 (surfaces in z = 0 are not supported)"};
      const size_t o = (size_t)g * n_panels + j;
      panels[o] = panel_record(proj, normal, centroid, area);
      if (inner) {
        panels[o].c[2] = 0.0;
        panels[o].interior = (double)raw.interior[j];
        panels[o].lid_eps = raw.eps.empty() ? 0.0 : raw.eps[j];
      }
      const int bi = raw.body[j];
      if (bi < 0 || bi >= n_bodies) throw Error{HBG_ERR_VALUE, "panel " + std::to_string(j + 1) + ": body out of range"};
      b[o] = bi;
      if (g >= grp.Gs) {  // a wall image moves as the mirror image of its real panel (lo_geometry.rs:981)
        const size_t r = (size_t)(g % grp.Gs) * n_panels + j;
        for (int d = 0; d < 6; ++d) nj[6 * o + d] = nj[6 * r + d];
        continue;
      }
      V3 rc = v3(raw.centres[3 * bi], raw.centres[3 * bi + 1], raw.centres[3 * bi + 2]);
      V3 mom = cross(centroid - rc, normal);
      for (int d = 0; d < 3; ++d) {  // the rigid modes do not move the interior free surface
        nj[6 * o + d] = inner ? 0.0 : normal[d];
        nj[6 * o + 3 + d] = inner ? 0.0 : mom[d];
      }
    }
  use_device(s);
  // from here the old geometry is gone: a failing upload leaves a session without panels
  s->statics_valid = false;
  s->stinf_valid = false;
  s->stf_valid = false;
  s->st32_valid = false;
  s->near_valid = false;
  s->field_F = -1;
  s->kochin_F = -1;
  s->N = 0;
  s->nb = 0;
  s->sym_x = sym_x;
  s->sym_y = sym_y;
  s->walls = walls;
  s->G = grp.G;
  s->Gs = grp.Gs;
  s->refl = grp.refl;
  s->chi = grp.chi;
  s->chisym = grp.chisym;
  s->d_chi.ensure(s->chi.size());
  HBG_HIP(hipMemcpy(s->d_chi.p, s->chi.data(), s->chi.size() * sizeof(double), hipMemcpyHostToDevice));
  s->d_chisym.ensure(s->chisym.size());
  HBG_HIP(hipMemcpy(s->d_chisym.p, s->chisym.data(), s->chisym.size() * sizeof(double), hipMemcpyHostToDevice));
  if (n_panels > 0) {
    s->d_panels.ensure(panels.size());
    s->d_body.ensure(b.size());
    s->d_nj.ensure(nj.size());
    HBG_HIP(hipMemcpy(s->d_panels.p, panels.data(), panels.size() * sizeof(PanelD), hipMemcpyHostToDevice));
    HBG_HIP(hipMemcpy(s->d_body.p, b.data(), b.size() * sizeof(int), hipMemcpyHostToDevice));
    HBG_HIP(hipMemcpy(s->d_nj.p, nj.data(), nj.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  s->panels = std::move(panels);
  s->body = std::move(b);
  s->nj = std::move(nj);
  s->raw_vertices = raw.vertices;
  s->raw_centres = raw.centres;
  s->raw_body = raw.body;
  s->raw_interior = raw.interior;
  s->raw_eps = raw.eps;
  s->NL = (int)std::count(raw.interior.begin(), raw.interior.end(), (char)2);
  s->N = n_panels;
  s->nb = n_bodies;
  s->NI = raw.NI;
}

// The panels in z = 0 after the body's: those of one kind (1 interior free surface, 2 lids) replaced
// by the given ones, those of the other kind kept. eps: the lids' damping factors (NULL: 0).
static HbgStatus set_sheet(HbgSession* s, char kind, const double* vertices, int32_t n_panels, const int32_t* body,
                           const double* eps, const char* where) {
  if (!s || (n_panels > 0 && !vertices)) return fail(HBG_ERR_ARGUMENT, "null argument", where);
  if (n_panels < 0) return fail(HBG_ERR_ARGUMENT, "negative count", where);
  if (s->N == 0) return fail(HBG_ERR_STATE, "no panels set", where);
  GUARD_BEGIN
  const char* what = kind == 1 ? "interior panel " : "lid panel ";
  const RawPanels raw = raw_of(s);
  const int nbody = raw.N - raw.NI;
  RawPanels out = raw;
  out.vertices.resize(12 * (size_t)nbody);
  out.body.resize(nbody);
  out.interior.assign(nbody, 0);
  out.eps.assign(nbody, 0.0);
  for (int j = nbody; j < raw.N; ++j)
    if (raw.interior[j] != kind) {
      out.vertices.insert(out.vertices.end(), raw.vertices.begin() + 12 * (size_t)j,
                          raw.vertices.begin() + 12 * (size_t)j + 12);
      out.body.push_back(raw.body[j]);
      out.interior.push_back(raw.interior[j]);
      out.eps.push_back(raw.eps.empty() ? 0.0 : raw.eps[j]);
    }
  for (int j = 0; j < n_panels; ++j) {
    for (int q = 0; q < 4; ++q)
      if (std::fabs(vertices[12 * (size_t)j + 3 * q + 2]) > 1e-9)
        throw Error{HBG_ERR_VALUE, what + std::to_string(j + 1) + ": a vertex off z = 0"};
    const double e = eps ? eps[j] : 0.0;
    if (!(e >= 0.0) || !std::isfinite(e))
      throw Error{HBG_ERR_VALUE, what + std::to_string(j + 1) + ": damping factor must be finite and >= 0"};
    out.vertices.insert(out.vertices.end(), vertices + 12 * (size_t)j, vertices + 12 * (size_t)j + 12);
    out.body.push_back(body ? body[j] : 0);
    out.interior.push_back(kind);
    out.eps.push_back(e);
  }
  out.N = (int)out.body.size();
  out.NI = out.N - nbody;
  set_geometry(s, make_group(s->sym_x, s->sym_y, s->walls), s->sym_x, s->sym_y, s->walls, out);
  return HBG_OK;
  GUARD_END(where)
}

HbgStatus hbg_set_interior(HbgSession* s, const double* vertices, int32_t n_panels, const int32_t* body) {
  return set_sheet(s, 1, vertices, n_panels, body, nullptr, "hbg_set_interior");
}

HbgStatus hbg_set_lids(HbgSession* s, const double* vertices, int32_t n_panels, const int32_t* body,
                       const double* damping_factor) {
  return set_sheet(s, 2, vertices, n_panels, body, damping_factor, "hbg_set_lids");
}

HbgStatus hbg_set_symmetry(HbgSession* s, int32_t sym_x, int32_t sym_y) {
  if (!s) return fail(HBG_ERR_ARGUMENT, "session is null", "hbg_set_symmetry");
  if ((sym_x != 0 && sym_x != 1) || (sym_y != 0 && sym_y != 1))
    return fail(HBG_ERR_VALUE, "sym_x and sym_y must be 0 or 1", "hbg_set_symmetry");
  if ((sym_x && (s->walls & 1)) || (sym_y && (s->walls & 2)))
    return fail(HBG_ERR_VALUE, "a symmetry plane where a wall stands", "hbg_set_symmetry");
  GUARD_BEGIN
  // panels that fail as images leave the planes, the walls and the geometry as they were
  set_geometry(s, make_group(sym_x, sym_y, s->walls), sym_x, sym_y, s->walls, raw_of(s));
  return HBG_OK;
  GUARD_END("hbg_set_symmetry")
}

HbgStatus hbg_set_field_points(HbgSession* s, const double* points, int32_t n) {
  if (!s || (n > 0 && !points)) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_set_field_points");
  if (n < 0) return fail(HBG_ERR_ARGUMENT, "negative count", "hbg_set_field_points");
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(points[3 * i]) || !std::isfinite(points[3 * i + 1]) || !(points[3 * i + 2] <= 0.0))
      return fail(HBG_ERR_VALUE, "field point " + std::to_string(i) + ": not finite or above z = 0",
                  "hbg_set_field_points");
  GUARD_BEGIN
  use_device(s);
  s->np = n;
  s->fpts.assign(points, points + 3 * (size_t)n);
  s->stf_valid = false;
  s->field_F = -1;
  if (n > 0) {
    std::vector<PanelD> rec(n);
    std::memset(rec.data(), 0, rec.size() * sizeof(PanelD));
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) rec[i].c[d] = points[3 * i + d];
    s->d_fpts.ensure(n);
    HBG_HIP(hipMemcpy(s->d_fpts.p, rec.data(), n * sizeof(PanelD), hipMemcpyHostToDevice));
  }
  return HBG_OK;
  GUARD_END("hbg_set_field_points")
}

HbgStatus hbg_field_values(HbgSession* s, double* frad, double* finc, double* fsca, double* ftot) {
  if (!s) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_field_values");
  if (s->field_F < 0) return fail(HBG_ERR_STATE, "no solve with field points yet", "hbg_field_values");
  const int F = s->field_F, P = s->np, K = s->field_nrad, H = s->field_H, W = K + 3 * H;
  for (int f = 0; f < F; ++f)
    for (int p = 0; p < P; ++p) {
      const double2* v = s->field_vals.data() + ((size_t)f * P + p) * W;
      for (int c = 0; c < K; ++c)
        if (frad) {
          frad[(((size_t)f * P + p) * K + c) * 2] = v[c].x;
          frad[(((size_t)f * P + p) * K + c) * 2 + 1] = v[c].y;
        }
      double* outs[3] = {finc, fsca, ftot};
      for (int q = 0; q < 3; ++q)
        if (outs[q])
          for (int h = 0; h < H; ++h) {
            outs[q][(((size_t)f * P + p) * H + h) * 2] = v[K + q * H + h].x;
            outs[q][(((size_t)f * P + p) * H + h) * 2 + 1] = v[K + q * H + h].y;
          }
    }
  return HBG_OK;
}

HbgStatus hbg_field_shape(HbgSession* s, int32_t* n_omegas, int32_t* n_points, int32_t* n_rad, int32_t* n_headings) {
  if (!s || !n_omegas || !n_points || !n_rad || !n_headings)
    return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_field_shape");
  if (s->field_F < 0) return fail(HBG_ERR_STATE, "no solve with field points yet", "hbg_field_shape");
  *n_omegas = s->field_F;
  *n_points = s->np;
  *n_rad = s->field_nrad;
  *n_headings = s->field_H;
  return HBG_OK;
}

HbgStatus hbg_kochin_shape(HbgSession* s, int32_t* n_omegas, int32_t* n_angles, int32_t* n_rad, int32_t* n_headings) {
  if (!s || !n_omegas || !n_angles || !n_rad || !n_headings)
    return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_kochin_shape");
  if (s->kochin_F < 0) return fail(HBG_ERR_STATE, "no solve with Kochin angles yet", "hbg_kochin_shape");
  *n_omegas = s->kochin_F;
  *n_angles = s->nk;
  *n_rad = s->kochin_nrad;
  *n_headings = s->kochin_H;
  return HBG_OK;
}

HbgStatus hbg_set_kochin(HbgSession* s, int32_t n_angles) {
  if (!s) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_set_kochin");
  if (n_angles < 0 || n_angles > 1 << 20) return fail(HBG_ERR_VALUE, "n_angles out of range", "hbg_set_kochin");
  s->nk = n_angles;
  s->kochin_F = -1;
  return HBG_OK;
}

HbgStatus hbg_kochin_values(HbgSession* s, double* kr, double* krt, double* kd, double* kdt) {
  if (!s) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_kochin_values");
  if (s->kochin_F < 0) return fail(HBG_ERR_STATE, "no solve with Kochin angles yet", "hbg_kochin_values");
  const int F = s->kochin_F, NK = s->nk, K = s->kochin_nrad, H = s->kochin_H, R = K + H;
  for (int f = 0; f < F; ++f)
    for (int d = 0; d < 2; ++d)
      for (int t = 0; t < NK; ++t) {
        const double2* v = s->kochin_vals.data() + (((size_t)f * 2 + d) * NK + t) * R;
        double* rad = d ? krt : kr;
        double* dif = d ? kdt : kd;
        if (rad)
          for (int c = 0; c < K; ++c) {
            rad[(((size_t)f * NK + t) * K + c) * 2] = v[c].x;
            rad[(((size_t)f * NK + t) * K + c) * 2 + 1] = v[c].y;
          }
        if (dif)
          for (int h = 0; h < H; ++h) {
            dif[(((size_t)f * NK + t) * H + h) * 2] = v[K + h].x;
            dif[(((size_t)f * NK + t) * H + h) * 2 + 1] = v[K + h].y;
          }
      }
  return HBG_OK;
}

HbgStatus hbg_set_walls(HbgSession* s, int32_t wall_x, int32_t wall_y) {
  if (!s) return fail(HBG_ERR_ARGUMENT, "session is null", "hbg_set_walls");
  if ((wall_x != 0 && wall_x != 1) || (wall_y != 0 && wall_y != 1))
    return fail(HBG_ERR_VALUE, "wall_x and wall_y must be 0 or 1", "hbg_set_walls");
  if ((wall_x && s->sym_x) || (wall_y && s->sym_y))
    return fail(HBG_ERR_VALUE, "a wall where a symmetry plane stands", "hbg_set_walls");
  GUARD_BEGIN
  // panels behind a new wall leave the walls and the geometry as they were
  const int walls = wall_x | wall_y << 1;
  set_geometry(s, make_group(s->sym_x, s->sym_y, walls), s->sym_x, s->sym_y, walls, raw_of(s));
  return HBG_OK;
  GUARD_END("hbg_set_walls")
}

// The group of the symmetry planes and the walls (lo_geometry.rs:863-945), its elements as bit
// masks (bit 0: x -> -x, bit 1: y -> -y): the symmetry elements first, then each of them times
// the wall reflections, so that the real panels (the sector and its images under the symmetry)
// come first and wall image block g mirrors real block g mod Gs. The characters of this
// product of reflections are chi_c(g) = (-1)^|c & g| (lo_geometry.rs:25-35 in the order e,
// sigma_x, sigma_y, sigma_x sigma_y); the solved ones, even under every wall reflection
// (lo_geometry.rs:1010-1030), are the Gs masks c of the symmetry: chis [Gs][G] for the assembly,
// chisym [Gs][Gs] over the symmetry elements for the right-hand sides and the recombination.
static Group make_group(int sym_x, int sym_y, int walls) {
  const int smask = sym_x | sym_y << 1, wmask = walls;
  std::vector<int> sym{0}, wal{0};
  for (int m = 1; m < 4; ++m) {
    if ((m & smask) == m) sym.push_back(m);
    if ((m & wmask) == m) wal.push_back(m);
  }
  std::vector<int> masks;
  for (int w : wal)
    for (int g : sym) masks.push_back(w | g);
  Group grp;
  grp.G = (int)masks.size();
  grp.Gs = (int)sym.size();
  grp.refl.clear();
  for (int m : masks) grp.refl.push_back({(m & 1) ? -1.0 : 1.0, (m & 2) ? -1.0 : 1.0});
  auto chi = [](int c, int g) { return __builtin_popcount(c & g) & 1 ? -1.0 : 1.0; };
  grp.chi.assign((size_t)grp.Gs * grp.G, 0.0);
  grp.chisym.assign((size_t)grp.Gs * grp.Gs, 0.0);
  for (int r = 0; r < grp.Gs; ++r) {
    for (int g = 0; g < grp.G; ++g) grp.chi[(size_t)r * grp.G + g] = chi(sym[r], masks[g]);
    for (int g = 0; g < grp.Gs; ++g) grp.chisym[(size_t)r * grp.Gs + g] = chi(sym[r], masks[g]);
  }
  return grp;
}

HbgStatus hbg_wave_kernel(HbgSession* s, const double* X, const double* Y, int64_t n, double* out) {
  if (!s || !X || !Y || !out) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_wave_kernel");
  if (n <= 0) return HBG_OK;
  for (int64_t i = 0; i < n; ++i)
    if (!(X[i] >= 0.0) || !(Y[i] >= 0.0) || (X[i] == 0.0 && Y[i] == 0.0) || !std::isfinite(X[i]) ||
        !std::isfinite(Y[i]))
      return fail(HBG_ERR_VALUE, "wave kernel point " + std::to_string(i) + " outside X, Y >= 0, (X, Y) != 0",
                  "hbg_wave_kernel");
  GUARD_BEGIN
  use_device(s);
  DevBuf<double> dx, dy, dout;
  dx.ensure(n);
  dy.ensure(n);
  dout.ensure(7 * (size_t)n);
  HBG_HIP(hipMemcpyAsync(dx.p, X, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
  HBG_HIP(hipMemcpyAsync(dy.p, Y, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
  launch_wave_points(s->T, dx.p, dy.p, (long)n, dout.p, s->stream);
  HBG_HIP(hipMemcpyAsync(out, dout.p, 7 * n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  dx.release();
  dy.release();
  dout.release();
  return HBG_OK;
  GUARD_END("hbg_wave_kernel")
}

HbgStatus hbg_finite_kernel(HbgSession* s, double h, double K, int32_t n, const double* R, const double* z,
                            const double* zeta, double* out) {
  if (!s || !R || !z || !zeta || !out) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_finite_kernel");
  if (!(h > 0.0) || !std::isfinite(h) || !(K >= 0.0) || n < 0)
    return fail(HBG_ERR_VALUE, "h must be finite and > 0, K >= 0 (0 and +infinity: the limits)", "hbg_finite_kernel");
  for (int i = 0; i < n; ++i)
    if (!(R[i] >= 0.0) || !(z[i] <= 0.0) || !(zeta[i] <= 0.0) || z[i] < -h || zeta[i] < -h ||
        (R[i] == 0.0 && z[i] + zeta[i] == 0.0))
      return fail(HBG_ERR_VALUE, "point " + std::to_string(i) + " outside R >= 0, -h <= z, zeta <= 0",
                  "hbg_finite_kernel");
  if (n == 0) return HBG_OK;
  GUARD_BEGIN
  use_device(s);
  if (K == 0.0 || std::isinf(K)) {
    // the frequency limits omega = 0 (K = 0) and omega = infinity (green_fin.rs FreqMode::Zero /
    // Infinite)
    DevBuf<double> dR, dz, dzeta, dout;
    dR.ensure(n);
    dz.ensure(n);
    dzeta.ensure(n);
    dout.ensure(8 * (size_t)n);
    HBG_HIP(hipMemcpy(dR.p, R, n * sizeof(double), hipMemcpyHostToDevice));
    HBG_HIP(hipMemcpy(dz.p, z, n * sizeof(double), hipMemcpyHostToDevice));
    HBG_HIP(hipMemcpy(dzeta.p, zeta, n * sizeof(double), hipMemcpyHostToDevice));
    launch_limit_points(h, K == 0.0, s->ktab, n, dR.p, dz.p, dzeta.p, dout.p, s->stream);
    HBG_HIP(hipMemcpyAsync(out, dout.p, 8 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HBG_HIP(hipStreamSynchronize(s->stream));
    return HBG_OK;
  }
  const FinFreq f = make_fin_freq(h, K);
  // the point depths as panels: z[i] at i, zeta[i] at n + i
  std::vector<PanelD> pts(2 * (size_t)n);
  std::memset(pts.data(), 0, pts.size() * sizeof(PanelD));
  for (int i = 0; i < n; ++i) {
    pts[i].c[2] = z[i];
    pts[n + i].c[2] = zeta[i];
  }
  DevBuf<PanelD> dp;
  DevBuf<double> dphi1, dphi2, dR, dz, dzeta, dout;
  DevBuf<double4> dpf;
  DevBuf<double2> dpcs;
  dp.ensure(pts.size());
  dphi1.ensure((size_t)FTS * FTW * FTILE);
  dphi2.ensure((size_t)FTS * FTV * FTILE);
  dpf.ensure(pts.size());
  dpcs.ensure(pts.size() * FIN_ROOTS);
  dR.ensure(n);
  dz.ensure(n);
  dzeta.ensure(n);
  dout.ensure(8 * (size_t)n);
  HBG_HIP(hipMemcpy(dp.p, pts.data(), pts.size() * sizeof(PanelD), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dR.p, R, n * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dz.p, z, n * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dzeta.p, zeta, n * sizeof(double), hipMemcpyHostToDevice));
  s->d_fwork.ensure(fin_tables_work());
  launch_fin_tables(f, s->T, s->d_gl.p + 16 * 32, dphi1.p, dphi2.p, s->fitm, dp.p, 2 * n, dpf.p, dpcs.p,
                    s->d_fwork.p, s->stream);
  FinTables F;
  F.phi1 = dphi1.p;
  F.phi2 = dphi2.p;
  F.ktab = s->ktab;
  F.pf = dpf.p;
  F.pcs = dpcs.p;
  DevBuf<FinFreq> dff;
  dff.ensure(1);
  HBG_HIP(hipMemcpy(dff.p, &f, sizeof(FinFreq), hipMemcpyHostToDevice));
  launch_fin_points_eval(dff.p, F, s->T, dR.p, dz.p, dzeta.p, n, dout.p, s->stream);
  HBG_HIP(hipMemcpyAsync(out, dout.p, 8 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  dp.release();
  dphi1.release();
  dphi2.release();
  dR.release();
  dz.release();
  dzeta.release();
  dout.release();
  dpf.release();
  dpcs.release();
  dff.release();
  return HBG_OK;
  GUARD_END("hbg_finite_kernel")
}

HbgStatus hbg_wave_kernel32(HbgSession* s, const double* X, const double* Y, int32_t n, double* out) {
  if (!s || !X || !Y || !out) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_wave_kernel32");
  if (n <= 0) return HBG_OK;
  for (int i = 0; i < n; ++i)
    if (!(X[i] >= 0.0) || !(Y[i] > 0.0) || !(X[i] < XT) || !(Y[i] < YT))
      return fail(HBG_ERR_VALUE, "point " + std::to_string(i) + " outside 0 <= X < 48, 0 < Y < 48",
                  "hbg_wave_kernel32");
  GUARD_BEGIN
  use_device(s);
  TablesOnDevice& t = tables32_for(s->device, s->stream);
  DevBuf<double> dx, dy, dout;
  dx.ensure(n);
  dy.ensure(n);
  dout.ensure(4 * (size_t)n);
  HBG_HIP(hipMemcpy(dx.p, X, n * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dy.p, Y, n * sizeof(double), hipMemcpyHostToDevice));
  launch_wave_points32(t.t2d32, t.t1d32, dx.p, dy.p, n, dout.p, s->stream);
  HBG_HIP(hipMemcpyAsync(out, dout.p, 4 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  return HBG_OK;
  GUARD_END("hbg_wave_kernel32")
}

HbgStatus hbg_finite_kernel32(HbgSession* s, double h, double K, int32_t n, const double* R, const double* z,
                              const double* zeta, double* out) {
  if (!s || !R || !z || !zeta || !out) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_finite_kernel32");
  if (!(h > 0.0) || !std::isfinite(h) || !(K > 0.0) || !std::isfinite(K) || n < 0)
    return fail(HBG_ERR_VALUE, "h and K must be finite and > 0", "hbg_finite_kernel32");
  const FinFreq f0 = make_fin_freq(h, K);
  for (int i = 0; i < n; ++i) {
    const bool nearf = R[i] < 0.5 * h;
    if (!(R[i] >= 0.0) || !(z[i] < 0.0) || !(zeta[i] < 0.0) || z[i] < -h || zeta[i] < -h ||
        (nearf && (K * R[i] >= XT || -K * (z[i] + zeta[i]) >= YT)) || f0.k0 * R[i] >= XT)
      return fail(HBG_ERR_VALUE, "point " + std::to_string(i) + " outside the FP32 tables or the water",
                  "hbg_finite_kernel32");
  }
  if (n == 0) return HBG_OK;
  GUARD_BEGIN
  use_device(s);
  TablesOnDevice& t = tables32_for(s->device, s->stream);
  std::vector<PanelD> pts(2 * (size_t)n);
  std::memset(pts.data(), 0, pts.size() * sizeof(PanelD));
  for (int i = 0; i < n; ++i) {
    pts[i].c[2] = z[i];
    pts[n + i].c[2] = zeta[i];
  }
  DevBuf<PanelD> dp;
  DevBuf<double> dphi1, dphi2, dR, dz, dzeta, dout;
  DevBuf<double4> dpf;
  DevBuf<double2> dpcs;
  DevBuf<float> p1, p2;
  DevBuf<float4> pf32;
  DevBuf<float2> pcs32;
  DevBuf<unsigned char> f32;
  dp.ensure(pts.size());
  dphi1.ensure((size_t)FTS * FTW * FTILE);
  dphi2.ensure((size_t)FTS * FTV * FTILE);
  dpf.ensure(pts.size());
  dpcs.ensure(pts.size() * FIN_ROOTS);
  p1.ensure(fin32_phi1_elems());
  p2.ensure(fin32_phi2_elems());
  pf32.ensure(pts.size());
  pcs32.ensure(pts.size() * FIN_ROOTS);
  f32.ensure(finfreq32_bytes());
  dR.ensure(n);
  dz.ensure(n);
  dzeta.ensure(n);
  dout.ensure(8 * (size_t)n);
  std::vector<unsigned char> hf(finfreq32_bytes());
  make_finfreq32(f0, hf.data());
  HBG_HIP(hipMemcpy(f32.p, hf.data(), hf.size(), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dp.p, pts.data(), pts.size() * sizeof(PanelD), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dR.p, R, n * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dz.p, z, n * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dzeta.p, zeta, n * sizeof(double), hipMemcpyHostToDevice));
  s->d_fwork.ensure(fin_tables_work());
  launch_fin_tables(f0, s->T, s->d_gl.p + 16 * 32, dphi1.p, dphi2.p, s->fitm, dp.p, 2 * n, dpf.p, dpcs.p,
                    s->d_fwork.p, s->stream);
  launch_fin_tables32(dphi1.p, dphi2.p, t.fin32P, p1.p, p2.p, dpf.p, dpcs.p, 2 * n, pf32.p, pcs32.p, s->stream);
  launch_fin_points32(f32.p, p1.p, p2.p, t.ktab32, pf32.p, pcs32.p, t.t2d32, t.t1d32, dR.p, dz.p, dzeta.p, n, dout.p,
                      s->stream);
  HBG_HIP(hipMemcpyAsync(out, dout.p, 8 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  return HBG_OK;
  GUARD_END("hbg_finite_kernel32")
}

HbgStatus hbg_dispersion_roots(double H, int32_t n, double* out) {
  if (!out || n < 0) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_dispersion_roots");
  if (!(H >= 0.0)) return fail(HBG_ERR_VALUE, "H must be >= 0", "hbg_dispersion_roots");
  for (int i = 0; i <= n; ++i) out[i] = dispersion_root(H, i);
  return HBG_OK;
}

HbgStatus hbg_statics(HbgSession* s, int32_t logarithmic, double log_tolerance, double* out) {
  if (!s || !out) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_statics");
  if (s->N == 0) return fail(HBG_ERR_STATE, "no panels set", "hbg_statics");
  if (s->G > 1) return fail(HBG_ERR_UNSUPPORTED, "the matrices of a body with symmetry planes", "hbg_statics");
  GUARD_BEGIN
  use_device(s);
  ensure_statics(s, logarithmic ? 1 : 0, log_tolerance);
  const int N = s->N;
  std::vector<StaticD> h((size_t)N * N);
  HBG_HIP(hipMemcpy(h.data(), s->d_st.p, h.size() * sizeof(StaticD), hipMemcpyDeviceToHost));
  for (int k = 0; k < N; ++k)
    for (int j = 0; j < N; ++j) {
      const StaticD& e = h[(size_t)j * N + k];
      double* o = out + ((size_t)k * N + j) * 4;
      o[0] = e.rs;
      o[1] = e.rd;
      o[2] = e.dl;
      o[3] = e.dd;
    }
  return HBG_OK;
  GUARD_END("hbg_statics")
}

HbgStatus hbg_assemble(HbgSession* s, double omega, const HbgOptions* options, double* S, double* A) {
  if (!s || !S || !A) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_assemble");
  if (s->N == 0) return fail(HBG_ERR_STATE, "no panels set", "hbg_assemble");
  if (s->G > 1) return fail(HBG_ERR_UNSUPPORTED, "the matrices of a body with symmetry planes", "hbg_assemble");
  if (!(omega > 0.0) || !std::isfinite(omega)) return fail(HBG_ERR_VALUE, "omega must be finite and > 0", "hbg_assemble");
  GUARD_BEGIN
  check_options(options);
  HbgOptions o = defaults(options);
  use_device(s);
  const bool m32 = o.precision == HBG_PRECISION_MIXED32;
  if (!m32) ensure_statics(s, o.logarithmic ? 1 : 0, o.log_tolerance);
  const int N = s->N;
  s->d_S.ensure((size_t)N * N);
  s->d_M.ensure((size_t)N * N);
  const double K = omega * omega / s->grav;
  check_depth(s);
  if (m32) {
    // the matrices of the FP32 assembly: its complex64 entries, the FP64 values at the near entries
    if (o.surrogate) throw Error{HBG_ERR_UNSUPPORTED, "the FP32 assembly with the surrogate"};
    if (s->NI > 0) throw Error{HBG_ERR_UNSUPPORTED, "the FP32 assembly with panels in z = 0 (an interior free surface or lids)"};
    if (!fp32_in_range(s, K)) throw Error{HBG_ERR_UNSUPPORTED, "pairs beyond the FP32 tables at this frequency"};
    ensure_fp32_geometry(s, o.logarithmic ? 1 : 0, o.log_tolerance);
    if (std::isfinite(s->depth)) {
      const FinFreq f = make_fin_freq(s->depth, K);
      std::vector<unsigned char> f32(finfreq32_bytes());
      make_finfreq32(f, f32.data());
      s->d_ffs.ensure(1);
      s->d_ffs32.ensure(f32.size());
      HBG_HIP(hipMemcpy(s->d_ffs.p, &f, sizeof(FinFreq), hipMemcpyHostToDevice));
      HBG_HIP(hipMemcpy(s->d_ffs32.p, f32.data(), f32.size(), hipMemcpyHostToDevice));
    }
    const size_t nnz = (size_t)std::max(s->near_nnz, 1);
    s->d_S32.ensure((size_t)N * N);
    s->d_dS.ensure(nnz);
    DevBuf<float2> a32;
    DevBuf<double2> dA;
    a32.ensure((size_t)N * N);
    dA.ensure(nnz);
    DevBuf<float2> a32r;
    a32r.ensure((size_t)N * N);
    assemble_frequency32(s, K, o.logarithmic ? 1.0 : 0.0, 0, a32.p, a32r.p, dA.p);
    std::vector<float2> hs((size_t)N * N), ha((size_t)N * N);
    std::vector<double2> hds(nnz), hda(nnz);
    HBG_HIP(hipMemcpyAsync(hs.data(), s->d_S32.p, hs.size() * sizeof(float2), hipMemcpyDeviceToHost, s->stream));
    HBG_HIP(hipMemcpyAsync(ha.data(), a32.p, ha.size() * sizeof(float2), hipMemcpyDeviceToHost, s->stream));
    HBG_HIP(hipMemcpyAsync(hds.data(), s->d_dS.p, nnz * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
    HBG_HIP(hipMemcpyAsync(hda.data(), dA.p, nnz * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
    HBG_HIP(hipStreamSynchronize(s->stream));
    for (int k = 0; k < N; ++k)
      for (int n = 0; n < N; ++n) {
        const size_t src = (size_t)n * N + k, dst = 2 * ((size_t)k * N + n);
        S[dst] = hs[src].x;
        S[dst + 1] = hs[src].y;
        A[dst] = ha[src].x;
        A[dst + 1] = ha[src].y;
      }
    for (int e = 0; e < s->near_nnz; ++e) {
      const size_t dst = 2 * ((size_t)s->near_ent[e].x * N + s->near_ent[e].y);
      S[dst] += hds[e].x;
      S[dst + 1] += hds[e].y;
      A[dst] += hda[e].x;
      A[dst + 1] += hda[e].y;
    }
    return HBG_OK;
  }
  assemble_frequency(s, K, o.logarithmic ? 1.0 : 0.0, s->d_M.p, o.surrogate != 0);
  std::vector<double2> hs((size_t)N * N), ha((size_t)N * N);
  HBG_HIP(hipMemcpyAsync(hs.data(), s->d_S.p, hs.size() * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
  HBG_HIP(hipMemcpyAsync(ha.data(), s->d_M.p, ha.size() * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  for (int k = 0; k < N; ++k)
    for (int n = 0; n < N; ++n) {
      const size_t src = (size_t)n * N + k, dst = 2 * ((size_t)k * N + n);
      S[dst] = hs[src].x;
      S[dst + 1] = hs[src].y;
      A[dst] = ha[src].x;
      A[dst + 1] = ha[src].y;
    }
  return HBG_OK;
  GUARD_END("hbg_assemble")
}

HbgStatus hbg_lu_solve(HbgSession* s, int32_t n, double* a, int32_t nrhs, double* b) {
  if (!s || !a || (nrhs > 0 && !b)) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_lu_solve");
  if (n <= 0 || nrhs < 0) return fail(HBG_ERR_ARGUMENT, "n must be > 0 and nrhs >= 0", "hbg_lu_solve");
  for (size_t i = 0; i < 2 * (size_t)n * n; ++i)
    if (!std::isfinite(a[i])) return fail(HBG_ERR_NUMERICAL, "non-finite system matrix", "hbg_lu_solve");
  GUARD_BEGIN
  use_device(s);
  DevBuf<double2> m;
  m.ensure((size_t)n * (n + nrhs));
  HBG_HIP(hipMemcpyAsync(m.p, a, (size_t)n * n * sizeof(double2), hipMemcpyHostToDevice, s->stream));
  if (nrhs > 0)
    HBG_HIP(hipMemcpyAsync(m.p + (size_t)n * n, b, (size_t)n * nrhs * sizeof(double2), hipMemcpyHostToDevice,
                           s->stream));
  const int zp = lu_solve_device(m.p, n, n, nrhs, s->stream, nullptr);
  HBG_HIP(hipMemcpyAsync(a, m.p, (size_t)n * n * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
  if (nrhs > 0)
    HBG_HIP(hipMemcpyAsync(b, m.p + (size_t)n * n, (size_t)n * nrhs * sizeof(double2), hipMemcpyDeviceToHost,
                           s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  m.release();
  if (zp > 0) throw Error{HBG_ERR_NUMERICAL, "singular system matrix (zero pivot " + std::to_string(zp) + ")"};
  for (size_t i = 0; i < 2 * (size_t)n * nrhs; ++i)
    if (!std::isfinite(b[i])) throw Error{HBG_ERR_NUMERICAL, "non-finite solution"};
  return HBG_OK;
  GUARD_END("hbg_lu_solve")
}

// The bytes hbg_lu_solve_noalloc needs in the scratch array: [A | B] (aligned), then the workspace.
static size_t noalloc_bytes(int n, int nrhs, size_t* matrix_bytes) {
  const size_t m = ((size_t)n * (n + nrhs) * sizeof(double2) + 255) / 256 * 256;
  if (matrix_bytes) *matrix_bytes = m;
  return m + lu_region_workspace_bytes(n);
}

HbgStatus hbg_lu_solve_noalloc(HbgSession* s, int32_t n, double* a, int32_t nrhs, double* b) {
  if (!s || !a || (nrhs > 0 && !b)) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_lu_solve_noalloc");
  if (n <= 0 || nrhs < 0) return fail(HBG_ERR_ARGUMENT, "n must be > 0 and nrhs >= 0", "hbg_lu_solve_noalloc");
  if (n > 1 << 16 || nrhs > 1 << 16 || noalloc_bytes(n, nrhs, nullptr) > noalloc_scratch_bytes())
    return fail(HBG_ERR_VALUE,
                "order " + std::to_string(n) + " with " + std::to_string(nrhs) +
                    " right-hand sides exceeds the allocation-free maximum (32 MiB of system and workspace)",
                "hbg_lu_solve_noalloc");
  for (size_t i = 0; i < 2 * (size_t)n * n; ++i)
    if (!std::isfinite(a[i])) return fail(HBG_ERR_NUMERICAL, "non-finite system matrix", "hbg_lu_solve_noalloc");
  GUARD_BEGIN
  use_device(s);
  size_t mb = 0;
  noalloc_bytes(n, nrhs, &mb);
  char* base = (char*)s->scratch;  // the device address taken at session creation
  double2* M = (double2*)base;
  auto upload = [&]() {
    HBG_HIP(hipMemcpyAsync(M, a, (size_t)n * n * sizeof(double2), hipMemcpyHostToDevice, s->stream));
    if (nrhs > 0)
      HBG_HIP(hipMemcpyAsync(M + (size_t)n * n, b, (size_t)n * nrhs * sizeof(double2), hipMemcpyHostToDevice,
                             s->stream));
  };
  upload();
  int zp = lu_solve_region(M, n, n, nrhs, base + mb, noalloc_scratch_bytes() - mb, 0, s->stream);
  if (zp < 0) {  // a panel needs interchanges: the system again from the caller's arrays, pivoted
    upload();
    zp = lu_solve_region(M, n, n, nrhs, base + mb, noalloc_scratch_bytes() - mb, 1, s->stream);
  }
  HBG_HIP(hipMemcpyAsync(a, M, (size_t)n * n * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
  if (nrhs > 0)
    HBG_HIP(hipMemcpyAsync(b, M + (size_t)n * n, (size_t)n * nrhs * sizeof(double2), hipMemcpyDeviceToHost,
                           s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  if (zp > 0) throw Error{HBG_ERR_NUMERICAL, "singular system matrix (zero pivot " + std::to_string(zp) + ")"};
  for (size_t i = 0; i < 2 * (size_t)n * nrhs; ++i)
    if (!std::isfinite(b[i])) throw Error{HBG_ERR_NUMERICAL, "non-finite solution"};
  return HBG_OK;
  GUARD_END("hbg_lu_solve_noalloc")
}

namespace {

// A frequency limit of the radiation problem (lo_frequency.rs with FreqMode::Zero / Infinite; no
// diffraction there): the matrices of the statics alone in infinite depth (launch_limit_matrices),
// in finite depth the statics and the limit's regular part (launch_limit_fin_matrices), S U for
// the radiation columns, the FP64 LU and the forces with omega = 0 in the force kernel (B = 0).
// A symmetric body or walls as at a frequency: the image blocks combined into the character
// systems, the right-hand sides projected on the characters, the potential recombined on the
// whole body (walls: the one even character, the forces on the real panels).
// Writes the added mass [D][K] into A (host). The statics (and at infinity their omega = infinity
// row) must be current.
void solve_limit(HbgSession* s, bool at_infinity, double* A) {
  const int N = s->N, nrad = 6 * s->nb, R = nrad, G = s->G, Gs = s->Gs, NF = Gs * N;
  const size_t nAB = (size_t)nrad * nrad, mstride = (size_t)N * (N + R);
  s->d_S.ensure((size_t)N * N * Gs);
  s->d_M.ensure(mstride * Gs);
  s->d_u.ensure((size_t)R * NF);
  s->d_part.ensure((size_t)rhs_chunks(N) * R * N * Gs);
  s->d_out.ensure(2 * nAB);
  s->d_phiI.ensure((size_t)NF);
  s->d_dphiI.ensure((size_t)NF);
  // the matrices: straight into S and M (G = 1), or the G image blocks
  double2* Sb = s->d_S.p;
  double2* Ab = s->d_M.p;
  if (G > 1) {
    s->d_Sg.ensure((size_t)N * N * G);
    s->d_Ag.ensure((size_t)N * N * G);
    Sb = s->d_Sg.p;
    Ab = s->d_Ag.p;
  }
  if (std::isfinite(s->depth))
    launch_limit_fin_matrices(s->d_panels.p, N, G * N, s->depth, !at_infinity, s->ktab, s->d_st.p, s->d_stinf.p, Sb,
                              Ab, s->stream);
  else
    launch_limit_matrices(s->d_st.p, at_infinity ? s->d_stinf.p : nullptr, N, G * N, Sb, Ab, s->stream);
  if (G > 1)
    launch_symmetry_combine(s->d_Sg.p, s->d_Ag.p, N, G, s->d_chi.p, Gs, s->d_S.p, s->d_M.p, mstride, N, nullptr,
                            s->stream);
  if (s->NI > 0) launch_interior_identity(s->d_panels.p, N, Gs, s->d_S.p, s->d_M.p, mstride, s->stream);
  launch_boundary(s->d_panels.p, s->d_body.p, s->d_nj.p, NF, nrad, 0, 1.0, 1.0, s->grav, 0.0, s->d_heads.p,
                  s->d_phiI.p, s->d_dphiI.p, s->d_u.p, s->stream, s->walls);
  if (Gs == 1) {
    launch_rhs(s->d_S.p, N, N, s->d_u.p, R, s->d_M.p, N, s->d_part.p, s->stream);
  } else {
    s->d_uc.ensure((size_t)N * R * Gs);
    launch_symmetry_project(s->d_u.p, N, Gs, R, s->d_chisym.p, s->d_uc.p, s->stream);
    launch_rhs_batched(s->d_S.p, (size_t)N * N, N, N, s->d_uc.p, (size_t)R * N, R, s->d_M.p, mstride, N, s->d_part.p,
                       Gs, s->stream);
  }
  for (int c = 0; c < Gs; ++c) {
    const int zp = lu_solve_device(s->d_M.p + mstride * c, N, N, R, s->stream, nullptr);
    if (zp > 0)
      throw Error{HBG_ERR_NUMERICAL, "singular system matrix at a frequency limit (zero pivot " + std::to_string(zp) +
                                         (Gs > 1 ? ", character " + std::to_string(c) : std::string()) + ")"};
  }
  if (Gs == 1) {
    launch_forces(s->d_panels.p, s->d_body.p, s->d_nj.p, N, nrad, 0, 0.0, s->rho, s->d_M.p, N, s->d_phiI.p,
                  s->d_dphiI.p, s->d_out.p, s->stream);
  } else {
    s->d_phif.ensure((size_t)NF * R);
    launch_symmetry_recombine(s->d_M.p + (size_t)N * N, mstride, N, N, Gs, R, s->d_chisym.p, s->d_phif.p, s->stream);
    launch_forces_sol(s->d_panels.p, s->d_body.p, s->d_nj.p, NF, nrad, 0, 0.0, s->rho, s->d_phif.p, NF, s->d_phiI.p,
                      s->d_dphiI.p, s->d_out.p, s->stream);
  }
  std::vector<double> h(2 * nAB);
  HBG_HIP(hipMemcpyAsync(h.data(), s->d_out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  for (double v : h)
    if (!std::isfinite(v)) throw Error{HBG_ERR_NUMERICAL, "non-finite solution at a frequency limit"};
  std::memcpy(A, h.data(), nAB * sizeof(double));
  if (std::isfinite(s->depth) && !at_infinity) {
    // omega = 0 in finite depth: a mode with a net volume flux through the body (heave of a
    // surface-piercing body) has an added mass that diverges as omega -> 0 and is not defined
    // (the reference refuses such a mode, rules/coordinates.rs:15-75, at the relative flux 1e-6):
    // its row and column are NaN. A mode whose normals vanish to round-off (yaw of a body of
    // revolution) carries no flux whatever its round-off sum.
    std::vector<double> flux(nrad, 0.0), scale(nrad, 0.0), body_scale(s->nb, 0.0);
    for (int j = 0; j < NF; ++j) {  // the whole body (its symmetric images; not a wall's)
      const PanelD& pj = s->panels[j];
      body_scale[s->body[j]] += pj.area * (1.0 + std::sqrt(pj.c[0] * pj.c[0] + pj.c[1] * pj.c[1] + pj.c[2] * pj.c[2]));
      for (int m = 0; m < 6; ++m) {
        const double f = s->nj[6 * (size_t)j + m] * pj.area;
        flux[6 * s->body[j] + m] += f;
        scale[6 * s->body[j] + m] += std::fabs(f);
      }
    }
    for (int c = 0; c < nrad; ++c)
      if (std::fabs(flux[c]) > 1e-6 * scale[c] && scale[c] > 1e-9 * body_scale[c / 6])
        for (int i = 0; i < nrad; ++i) A[(size_t)c * nrad + i] = A[(size_t)i * nrad + c] = NAN;
  }
}

}  // namespace

HbgStatus hbg_solve(HbgSession* s, const double* omegas, int32_t n_omegas, const double* headings_deg,
                    int32_t n_headings, const HbgOptions* options, double* added_mass, double* damping,
                    double* excitation, double* froude_krylov, double* scattering, double* haskind,
                    HbgProgressFn progress, void* user) {
  if (!s || (n_omegas > 0 && !omegas) || (n_headings > 0 && !headings_deg))
    return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_solve");
  if (n_omegas < 0 || n_headings < 0) return fail(HBG_ERR_ARGUMENT, "negative count", "hbg_solve");
  if (s->N == 0) return fail(HBG_ERR_STATE, "no panels set", "hbg_solve");
  for (int i = 0; i < n_omegas; ++i)
    if (!(omegas[i] >= 0.0))
      return fail(HBG_ERR_VALUE,
                  "omega " + std::to_string(i) + " must be > 0, or 0 or +infinity (the radiation limits)", "hbg_solve");
  GUARD_BEGIN
  check_options(options);
  HbgOptions o = defaults(options);
  use_device(s);
  const int N = s->N, H = n_headings, nrad = 6 * s->nb, D = nrad, R = nrad + H;
  if (R > 32) {
    // more headings than the 32 right-hand sides of one system: the headings in groups, each
    // solved with the radiation columns (the Haskind excitation of a heading takes them), the
    // outputs of each group into the caller's arrays (the matrices are assembled and factored
    // once per group)
    const int hmax = 32 - nrad;
    if (hmax < 1) throw Error{HBG_ERR_UNSUPPORTED, "more than 31 radiation columns (6 per body)"};
    if (s->np > 0 || s->nk > 0)
      throw Error{HBG_ERR_UNSUPPORTED, "field points or Kochin functions with more than 32 right-hand sides"};
    const size_t F = (size_t)n_omegas;
    std::vector<double> x, fk, sc, hk;
    for (int h0 = 0; h0 < H; h0 += hmax) {
      const int hc = std::min(hmax, H - h0);
      x.assign(F * hc * D * 2, 0.0);
      fk.assign(x.size(), 0.0);
      sc.assign(x.size(), 0.0);
      hk.assign(x.size(), 0.0);
      const HbgStatus st = hbg_solve(s, omegas, n_omegas, headings_deg + h0, hc, options, h0 ? nullptr : added_mass,
                                     h0 ? nullptr : damping, x.data(), fk.data(), sc.data(), hk.data(), progress, user);
      if (st != HBG_OK) return st;
      for (size_t f = 0; f < F; ++f)
        for (int hh = 0; hh < hc; ++hh)
          for (int d = 0; d < D; ++d)
            for (int c = 0; c < 2; ++c) {
              const size_t src = ((f * hc + hh) * D + d) * 2 + c, dst = ((f * H + h0 + hh) * D + d) * 2 + c;
              if (excitation) excitation[dst] = x[src];
              if (froude_krylov) froude_krylov[dst] = fk[src];
              if (scattering) scattering[dst] = sc[src];
              // the Haskind excitation [F][D][H]
              if (haskind) haskind[((f * D + d) * H + h0 + hh) * 2 + c] = hk[((f * D + d) * hc + hh) * 2 + c];
            }
    }
    return HBG_OK;
  }
  check_depth(s);
  // the ordinary frequencies go through the batched path (om[i] is the caller's omegas[io[i]]);
  // the limits omega = 0 and infinity, infinite depth only, are solved after them (solve_limit)
  std::vector<double> om;
  std::vector<int> io, il;
  bool want_inf = false;
  for (int i = 0; i < n_omegas; ++i) {
    if (omegas[i] > 0.0 && std::isfinite(omegas[i])) {
      om.push_back(omegas[i]);
      io.push_back(i);
    } else {
      il.push_back(i);
      want_inf = want_inf || std::isinf(omegas[i]);
    }
  }
  if (s->NI > 0 && o.surrogate) throw Error{HBG_ERR_UNSUPPORTED, "the surrogate with panels in z = 0 (an interior free surface or lids)"};
  if ((s->nk > 0 || s->np > 0) && !il.empty())
    throw Error{HBG_ERR_UNSUPPORTED, "field points or Kochin functions at the frequency limits"};
  if (s->np > 0 && s->NL > 0) throw Error{HBG_ERR_UNSUPPORTED, "field points with lids"};
  if (s->np > 0 && std::isfinite(s->depth))
    for (int i = 0; i < s->np; ++i)
      if (s->fpts[3 * i + 2] < -s->depth)
        throw Error{HBG_ERR_VALUE, "field point " + std::to_string(i) + " below the sea floor"};
  // a symmetric body (G > 1): every frequency is G character systems of order N (the sector),
  // consecutive in the batch buffers; the boundary values and the forces take the whole body.
  // With walls one system of order N (the even character), the boundary values and the forces on
  // the real panels, the wall images only in the matrices.
  const int G = s->Gs, NF = G * N;
  const int NS = s->G * N;  // the source panels of the field values and Kochin functions, images included
  const double* const omegas_all = omegas;
  omegas = om.data();
  n_omegas = (int32_t)om.size();
  if (progress && progress(user, HBG_STAGE_STATICS, 0, 1) == 0) throw Error{HBG_ERR_CANCELLED, "cancelled"};
  // the FP32 far-field assembly: one geometry without images, no surrogate, every frequency's
  // pairs inside the FP32 tables; otherwise HBG_PRECISION_MIXED. It takes its own statics (FP32
  // records, FP64 at the near entries); the FP64 statics of every pair only for the limits.
  bool m32 = o.precision == HBG_PRECISION_MIXED32 && s->G == 1 && !o.surrogate && s->NI == 0;
  for (int i = 0; m32 && i < n_omegas; ++i) m32 = fp32_in_range(s, omegas[i] * omegas[i] / s->grav);
  if (!m32 || !il.empty()) ensure_statics(s, o.logarithmic ? 1 : 0, o.log_tolerance, want_inf);
  if (m32) ensure_fp32_geometry(s, o.logarithmic ? 1 : 0, o.log_tolerance);
  if (s->np > 0) ensure_field_statics(s, o.logarithmic ? 1 : 0, o.log_tolerance);
  std::vector<double> heads(std::max(H, 1), 0.0);
  for (int h = 0; h < H; ++h) heads[h] = headings_deg[h] * PI / 180.0;
  s->d_heads.ensure(heads.size());
  HBG_HIP(hipMemcpy(s->d_heads.p, heads.data(), heads.size() * sizeof(double), hipMemcpyHostToDevice));
  // finite depth: every frequency's roots and coefficients on the device before the first launch,
  // so that no pageable copy (which waits for the stream) sits between the batches' launches
  if (std::isfinite(s->depth) && n_omegas > 0) {
    std::vector<FinFreq> ffs(n_omegas);
    for (int i = 0; i < n_omegas; ++i) ffs[i] = make_fin_freq(s->depth, omegas[i] * omegas[i] / s->grav);
    s->d_ffs.ensure(ffs.size());
    HBG_HIP(hipMemcpy(s->d_ffs.p, ffs.data(), ffs.size() * sizeof(FinFreq), hipMemcpyHostToDevice));
  }
  const size_t nAB = (size_t)D * nrad, nX = (size_t)H * D, nHK = (size_t)nrad * H;
  const size_t nout = 2 * nAB + 6 * nX + 2 * nHK;
  // frequencies in flight: each holds its [A | B] (N x (N + R)) and incident-wave values until
  // its forces are taken; S, the right-hand-side partials and the finite-depth tables are shared
  const size_t mstride = (size_t)N * (N + R);
  const size_t hstride = (size_t)std::max(H, 1) * NF;
  const bool mixed = o.precision == HBG_PRECISION_MIXED || (o.precision == HBG_PRECISION_MIXED32 && !m32);
  const bool prof = s->profiling;
  const double m32_tol = std::pow(10.0, -(double)g_mixed32_tol_exp);
  if (m32) {
    if (std::isfinite(s->depth) && n_omegas > 0) {
      std::vector<unsigned char> f32((size_t)n_omegas * finfreq32_bytes());
      for (int i = 0; i < n_omegas; ++i)
        make_finfreq32(make_fin_freq(s->depth, omegas[i] * omegas[i] / s->grav), f32.data() + i * finfreq32_bytes());
      s->d_ffs32.ensure(f32.size());
      HBG_HIP(hipMemcpy(s->d_ffs32.p, f32.data(), f32.size(), hipMemcpyHostToDevice));
    }
  }
  const size_t nnzc = (size_t)std::max(s->near_nnz, 1);  // near entries per system (FP32 assembly)
  const size_t bstride = (size_t)N * R;                   // FP64 right-hand sides per system (FP32 assembly)
  // two halves of the batch buffers: the assembly of batch t + 1 (session stream) runs beside the
  // LU, refinement and forces of batch t (LU stream). One half (and one stream) when one batch
  // takes every frequency, or with stage timing. The batch size is balanced over the batches.
  int B = batch_size(s, N, R, n_omegas, mixed || m32, G);
  const int halves = (prof || B >= n_omegas) ? 1 : 2;
  if (halves == 2) B = batch_size(s, N, R, n_omegas, mixed || m32, 2 * G);
  if (B > 0 && n_omegas > 0) {
    const int nbt = (n_omegas + B - 1) / B;
    B = (n_omegas + nbt - 1) / nbt;
  }
  hipStream_t sa = s->stream, sb = halves == 1 ? s->stream : s->lu_stream;
  const int BG = B * G;  // systems in flight per half
  s->d_S.ensure((size_t)N * N * G);
  // the FP32 assembly keeps no FP64 matrices in flight: one FP64 system for a frequency that
  // needs the FP64 path
  s->d_M.ensure(m32 ? mstride : mstride * BG * halves);
  if (m32) {
    s->d_M32.ensure(mstride * BG * halves);
    s->d_B64.ensure(bstride * BG * halves);
    s->d_S32.ensure((size_t)N * N);
    s->d_A32r.ensure((size_t)N * N * BG * halves);
    s->d_dS.ensure(nnzc);
    s->d_dA.ensure(nnzc * BG * halves);
  }
  s->d_phiI.ensure(hstride * B * halves);
  s->d_dphiI.ensure(hstride * B * halves);
  s->d_u.ensure((size_t)R * NF);
  s->d_part.ensure((size_t)rhs_chunks(N) * R * N * G);
  s->d_out.ensure(nout * B * halves);
  if (G > 1) {
    s->d_uc.ensure((size_t)R * N * G);
    s->d_phif.ensure((size_t)R * NF);
  }
  if (mixed || m32) {
    s->d_M32.ensure(mstride * BG * halves);
    s->d_D32.ensure((size_t)N * R * BG);
    s->d_X64.ensure((size_t)N * R * BG);
    s->d_rpart.ensure(residual_max_elems(N) * BG);
    s->d_rpart2.ensure(residual_part_elems(N, R) * BG);
    s->d_rscale.ensure((size_t)R * BG);
    s->d_active.ensure((size_t)BG);
  }
  std::vector<double> hout(nout * B);
  // field values: per frequency [np][nrad + 3 H] on the device, all frequencies on the host
  const size_t fw = (size_t)nrad + 3 * (size_t)H, fstride = (size_t)s->np * fw;
  const size_t kstride = 2 * (size_t)s->nk * R;
  std::vector<double2> hkochin;
  if (s->nk > 0) {
    s->d_kout.ensure(kstride * B * halves);
    hkochin.resize(kstride * B);
    s->kochin_vals.assign(kstride * (size_t)n_omegas, make_double2(0.0, 0.0));
    s->kochin_F = -1;
  }
  std::vector<double2> hfield;
  if (s->np > 0) {
    s->d_Sx.ensure((size_t)NS * s->np);
    s->d_Ax.ensure((size_t)NS * s->np);
    s->d_fout.ensure(fstride * B * halves);
    if (std::isfinite(s->depth)) {
      s->d_phi1f.ensure((size_t)FTS * FTW * FTILE);
      s->d_phi2f.ensure((size_t)FTS * FTV * FTILE);
      s->d_pff.ensure((size_t)NS + s->np);
      s->d_pcsf.ensure(((size_t)NS + s->np) * FIN_ROOTS);
      s->d_fworkf.ensure(fin_tables_work());
    }
    hfield.resize(fstride * B);
    s->field_vals.assign(fstride * (size_t)n_omegas, make_double2(0.0, 0.0));
    s->field_F = -1;
  }
  std::vector<int> iters(BG);
  std::vector<double> eta(BG);
  if (prof && s->bev.size() < (size_t)(3 * B + 3)) {
    for (size_t i = s->bev.size(); i < (size_t)(3 * B + 3); ++i) {
      hipEvent_t e;
      HBG_HIP(hipEventCreate(&e));
      s->bev.push_back(e);
    }
  }
  const double logc = o.logarithmic ? 1.0 : 0.0;
  const double depth = std::isfinite(s->depth) ? s->depth : 0.0;
  // assembly, incident wave and S U of frequency fi into slot b of the buffers' half h (session
  // stream); the slot's place in the buffers is slot(h, b)
  auto slot = [&](int h, int b) { return (size_t)h * B + b; };
  auto build = [&](int fi, int b, int h) {
    const double w = omegas[fi], K = w * w / s->grav;
    double2* Mb = s->d_M.p + mstride * slot(h, b) * G;
    if (prof) HBG_HIP(hipEventRecord(s->bev[3 * b], sa));
    if (m32) {
      // the FP32 assembly into the slot's complex64 system, the right-hand sides in FP64
      double2* Bb = s->d_B64.p + bstride * slot(h, b);
      const double kw = assemble_frequency32(s, K, logc, fi, s->d_M32.p + mstride * slot(h, b),
                                             s->d_A32r.p + (size_t)N * N * slot(h, b), s->d_dA.p + nnzc * slot(h, b));
      launch_boundary(s->d_panels.p, s->d_body.p, s->d_nj.p, NF, nrad, H, w, kw, s->grav, depth, s->d_heads.p,
                      s->d_phiI.p + hstride * slot(h, b), s->d_dphiI.p + hstride * slot(h, b), s->d_u.p, sa,
                      s->walls);
      if (prof) HBG_HIP(hipEventRecord(s->bev[3 * b + 1], sa));
      launch_rhs32(s->d_S32.p, N, s->d_u.p, R, Bb, s->d_part.p, s->d_near_rowptr.p, s->d_near_ent.p, s->d_dS.p, sa);
      if (prof) HBG_HIP(hipEventRecord(s->bev[3 * b + 2], sa));
      return;
    }
    const double kw = assemble_frequency(s, K, logc, Mb, o.surrogate != 0,
                                         mixed ? s->d_M32.p + mstride * slot(h, b) * G : nullptr, N, mstride,
                                         std::isfinite(s->depth) ? s->d_ffs.p + fi : nullptr);
    launch_boundary(s->d_panels.p, s->d_body.p, s->d_nj.p, NF, nrad, H, w, kw, s->grav, depth, s->d_heads.p,
                    s->d_phiI.p + hstride * slot(h, b), s->d_dphiI.p + hstride * slot(h, b), s->d_u.p, sa, s->walls);
    if (prof) HBG_HIP(hipEventRecord(s->bev[3 * b + 1], sa));
    if (G == 1) {
      launch_rhs(s->d_S.p, N, N, s->d_u.p, R, Mb, N, s->d_part.p, sa);
    } else {
      // every character c: S_c times the projection of u (lo_frequency.rs:873-893), one launch
      launch_symmetry_project(s->d_u.p, N, G, R, s->d_chisym.p, s->d_uc.p, sa);
      launch_rhs_batched(s->d_S.p, (size_t)N * N, N, N, s->d_uc.p, (size_t)R * N, R, Mb, mstride, N, s->d_part.p, G,
                         sa);
    }
    if (prof) HBG_HIP(hipEventRecord(s->bev[3 * b + 2], sa));
  };
  // a rebuilt system of the LU stream's batch: the LU stream waits for it
  auto rebuild = [&](int fi, int b, int h) {
    build(fi, b, h);
    if (sb != sa) {
      HBG_HIP(hipEventRecord(s->assembled[h], sa));
      HBG_HIP(hipStreamWaitEvent(sb, s->assembled[h], 0));
    }
  };
  // a system of the FP32 assembly that needs the FP64 path: FP64 assembly into the one FP64
  // system, its LU (pivoted where needed), the solution into the slot's right-hand sides
  auto fallback64 = [&](int fi, int b, int h) {
    const double w = omegas[fi], K = w * w / s->grav;
    // the FP64 statics of every pair (the FP32 assembly computed them at its near entries only),
    // on the session stream after the assembly work queued there
    ensure_statics(s, o.logarithmic ? 1 : 0, o.log_tolerance);
    const double kw = assemble_frequency(s, K, logc, s->d_M.p, false, nullptr, N, mstride,
                                         std::isfinite(s->depth) ? s->d_ffs.p + fi : nullptr);
    launch_boundary(s->d_panels.p, s->d_body.p, s->d_nj.p, NF, nrad, H, w, kw, s->grav, depth, s->d_heads.p,
                    s->d_phiI.p + hstride * slot(h, b), s->d_dphiI.p + hstride * slot(h, b), s->d_u.p, sa, s->walls);
    launch_rhs(s->d_S.p, N, N, s->d_u.p, R, s->d_M.p, N, s->d_part.p, sa);
    if (sb != sa) {
      HBG_HIP(hipEventRecord(s->assembled[h], sa));
      HBG_HIP(hipStreamWaitEvent(sb, s->assembled[h], 0));
    }
    const int zp = lu_solve_device(s->d_M.p, N, N, R, sb, nullptr);
    if (zp > 0) throw Error{HBG_ERR_NUMERICAL, "singular system matrix (zero pivot " + std::to_string(zp) + ")"};
    HBG_HIP(hipMemcpyAsync(s->d_B64.p + bstride * slot(h, b), s->d_M.p + (size_t)N * N, bstride * sizeof(double2),
                           hipMemcpyDeviceToDevice, sb));
    if (prof) s->prof[15] += 1.0;
  };
  const int nbatches = (n_omegas + B - 1) / B;
  auto enqueue_batch = [&](int t) {
    const int f0 = t * B, nb = std::min(B, n_omegas - f0), h = t % halves;
    for (int b = 0; b < nb; ++b) {
      if (progress && progress(user, HBG_STAGE_ASSEMBLY, f0 + b, n_omegas) == 0)
        throw Error{HBG_ERR_CANCELLED, "cancelled"};
      build(f0 + b, b, h);
    }
    if (sb != sa) HBG_HIP(hipEventRecord(s->assembled[h], sa));
  };
  auto stage_ms = [&](hipEvent_t a0, hipEvent_t a1) {
    float ms = 0.0f;
    HBG_HIP(hipEventSynchronize(a1));
    HBG_HIP(hipEventElapsedTime(&ms, a0, a1));
    return (double)ms;
  };
  std::vector<int> status(BG);
  if (nbatches > 0) enqueue_batch(0);
  for (int t = 0; t < nbatches; ++t) {
    const int f0 = t * B, nb = std::min(B, n_omegas - f0), h = t % halves, nbg = nb * G;
    double2* Mh = s->d_M.p + mstride * slot(h, 0) * G;
    // the next batch's assembly goes ahead on the session stream (its half was last used by
    // batch t - 1, whose results the host has already copied)
    if (halves > 1 && t + 1 < nbatches) enqueue_batch(t + 1);
    if (sb != sa) HBG_HIP(hipStreamWaitEvent(sb, s->assembled[h], 0));
    if (prof)
      for (int b = 0; b < nb; ++b) {
        s->prof[1] += stage_ms(s->bev[3 * b], s->bev[3 * b + 1]);
        s->prof[2] += stage_ms(s->bev[3 * b + 1], s->bev[3 * b + 2]);
      }
    for (int b = 0; b < nb; ++b)
      if (progress && progress(user, HBG_STAGE_LU, f0 + b, n_omegas) == 0) throw Error{HBG_ERR_CANCELLED, "cancelled"};
    LuStats lus;
    if (G > 1) {
      // the character systems of the batch; a frequency with a system that needs the FP64 or the
      // pivoted path is rebuilt (all its characters) and each solved by the FP64 LU with fallback
      if (mixed)
        mixed_solve_batched(Mh, s->d_M32.p + mstride * slot(h, 0) * G, s->d_X64.p, s->d_D32.p, s->d_rpart.p,
                            s->d_rpart2.p, s->d_rscale.p, nbg, N, R, true, MIXED_TOL, MIXED_MAX_ITER, sb,
                            prof ? &lus : nullptr, status.data(), iters.data(), eta.data(), nullptr, s->d_active.p);
      else
        lu_solve_batched(Mh, mstride, nbg, N, N, R, sb, prof ? &lus : nullptr, status.data());
      for (int b = 0; b < nb; ++b) {
        bool again = false;
        for (int c = 0; c < G; ++c) {
          const int q = b * G + c;
          if (status[q] > 0)
            throw Error{HBG_ERR_NUMERICAL, "singular system matrix (zero pivot " + std::to_string(status[q]) + ")"};
          again = again || status[q] < 0;
          if (mixed && prof && status[q] == 0) {
            s->prof[13] += iters[q];
            s->prof[14] = std::max(s->prof[14], eta[q]);
          }
        }
        if (!again) continue;
        rebuild(f0 + b, b, h);
        for (int c = 0; c < G; ++c) {
          const int zp = lu_solve_device(Mh + mstride * (b * G + c), N, N, R, sb, prof ? &lus : nullptr);
          if (zp > 0) throw Error{HBG_ERR_NUMERICAL, "singular system matrix (zero pivot " + std::to_string(zp) + ")"};
        }
        if (mixed && prof) s->prof[15] += 1.0;
      }
      std::fill(status.begin(), status.begin() + nbg, 0);
    } else if (m32) {
      mixed32_solve_batched(s->d_M32.p + mstride * slot(h, 0), s->d_A32r.p + (size_t)N * N * slot(h, 0),
                            s->d_B64.p + bstride * slot(h, 0),
                            s->d_near_rowptr.p, s->d_near_ent.p, s->d_dA.p + nnzc * slot(h, 0), nnzc, s->d_X64.p,
                            s->d_D32.p, s->d_rpart.p, s->d_rpart2.p, s->d_rscale.p, nb, N, R, m32_tol,
                            MIXED_MAX_ITER, sb, prof ? &lus : nullptr, status.data(), iters.data(), eta.data(),
                            s->d_active.p);
      for (int b = 0; b < nb; ++b) {
        if (status[b] > 0)
          throw Error{HBG_ERR_NUMERICAL, "singular system matrix (zero pivot " + std::to_string(status[b]) + ")"};
        if (status[b] < 0) {
          fallback64(f0 + b, b, h);
        } else if (prof) {
          s->prof[13] += iters[b];
          s->prof[14] = std::max(s->prof[14], eta[b]);
        }
        status[b] = 0;
      }
    } else if (mixed) {
      mixed_solve_batched(Mh, s->d_M32.p + mstride * slot(h, 0), s->d_X64.p, s->d_D32.p, s->d_rpart.p, s->d_rpart2.p,
                          s->d_rscale.p, nb, N, R, true, MIXED_TOL, MIXED_MAX_ITER, sb, prof ? &lus : nullptr,
                          status.data(), iters.data(), eta.data(), nullptr, s->d_active.p);
      for (int b = 0; b < nb; ++b) {
        if (status[b] < 0) {  // the plain FP64 path for this frequency
          rebuild(f0 + b, b, h);
          lu_solve_batched(Mh + mstride * b, mstride, 1, N, N, R, sb, prof ? &lus : nullptr, &status[b]);
          if (prof) s->prof[15] += 1.0;
        } else if (prof) {
          s->prof[13] += iters[b];
          s->prof[14] = std::max(s->prof[14], eta[b]);
        }
      }
    } else {
      lu_solve_batched(Mh, mstride, nb, N, N, R, sb, prof ? &lus : nullptr, status.data());
    }
    for (int b = 0; b < nb && G == 1; ++b) {
      if (status[b] < 0) {  // a panel needs interchanges: rebuild the system, pivoted path
        rebuild(f0 + b, b, h);
        status[b] = lu_solve_pivoted(Mh + mstride * b, N, N, R, sb, prof ? &lus : nullptr);
      }
      if (status[b] > 0)
        throw Error{HBG_ERR_NUMERICAL, "singular system matrix (zero pivot " + std::to_string(status[b]) + ")"};
    }
    const int e0 = 3 * B, e1 = 3 * B + 1;
    if (prof) HBG_HIP(hipEventRecord(s->bev[e0], sb));
    for (int b = 0; b < nb; ++b) {
      if (m32) {
        launch_forces_sol(s->d_panels.p, s->d_body.p, s->d_nj.p, N, nrad, H, omegas[f0 + b], s->rho,
                          s->d_B64.p + bstride * slot(h, b), N, s->d_phiI.p + hstride * slot(h, b),
                          s->d_dphiI.p + hstride * slot(h, b), s->d_out.p + nout * slot(h, b), sb);
      } else if (G == 1) {
        launch_forces(s->d_panels.p, s->d_body.p, s->d_nj.p, N, nrad, H, omegas[f0 + b], s->rho, Mh + mstride * b, N,
                      s->d_phiI.p + hstride * slot(h, b), s->d_dphiI.p + hstride * slot(h, b),
                      s->d_out.p + nout * slot(h, b), sb);
      } else {
        // the potential on the whole body from the characters (lo_frequency.rs:931-943)
        launch_symmetry_recombine(Mh + mstride * b * G + (size_t)N * N, mstride, N, N, G, R, s->d_chisym.p,
                                  s->d_phif.p, sb);
        launch_forces_sol(s->d_panels.p, s->d_body.p, s->d_nj.p, NF, nrad, H, omegas[f0 + b], s->rho, s->d_phif.p,
                          NF, s->d_phiI.p + hstride * slot(h, b), s->d_dphiI.p + hstride * slot(h, b),
                          s->d_out.p + nout * slot(h, b), sb);
      }
      // Kochin functions and field values from the potential on all NF panels (a symmetric body:
      // recombined just above, before the next frequency reuses the buffer)
      const double2* solb = m32 ? s->d_B64.p + bstride * slot(h, b)
                                : (G == 1 ? Mh + mstride * b + (size_t)N * N : s->d_phif.p);
      const double w = omegas[f0 + b], K = w * w / s->grav;
      if (s->nk > 0) {
        const double kw = std::isfinite(s->depth) ? make_fin_freq(s->depth, K).k0 : K;
        launch_kochin(s->d_panels.p, s->d_body.p, s->d_nj.p, NS, NF, nrad, H, kw,
                      std::isfinite(s->depth) ? s->depth : 0.0, s->nk, solb, s->d_phiI.p + hstride * slot(h, b),
                      s->d_kout.p + kstride * slot(h, b), sb);
      }
      if (s->np > 0) {
        double kw = K;
        FinTables Ff{};
        if (std::isfinite(s->depth)) {
          // the frequency's tables for the panels and the points, on the LU stream
          const FinFreq f = make_fin_freq(s->depth, K);
          kw = f.k0;
          launch_fin_tables(f, s->T, s->d_gl.p + 16 * 32, s->d_phi1f.p, s->d_phi2f.p, s->fitm, s->d_fcomb.p,
                            NS + s->np, s->d_pff.p, s->d_pcsf.p, s->d_fworkf.p, sb);
          Ff = FinTables{s->d_phi1f.p, s->d_phi2f.p, s->ktab, s->d_pff.p, s->d_pcsf.p};
        }
        // the tables of the next frequency overwrite these on the same stream, after this launch
        launch_field_values(s->d_fpts.p, s->d_panels.p, s->d_body.p, s->d_nj.p, s->np, NS, NF, nrad, H, w, K, kw,
                            std::isfinite(s->depth) ? s->depth : 0.0,
                            std::isfinite(s->depth) ? s->d_ffs.p + f0 + b : nullptr, Ff, logc, s->T, s->grav,
                            s->d_heads.p, s->walls, s->d_stf.p, s->d_Sx.p, s->d_Ax.p, solb,
                            s->d_phiI.p + hstride * slot(h, b), s->d_fout.p + fstride * slot(h, b), sb);
      }
    }
    if (prof) HBG_HIP(hipEventRecord(s->bev[e1], sb));
    if (s->nk > 0)
      HBG_HIP(hipMemcpyAsync(hkochin.data(), s->d_kout.p + kstride * slot(h, 0), nb * kstride * sizeof(double2),
                             hipMemcpyDeviceToHost, sb));
    if (s->np > 0)
      HBG_HIP(hipMemcpyAsync(hfield.data(), s->d_fout.p + fstride * slot(h, 0), nb * fstride * sizeof(double2),
                             hipMemcpyDeviceToHost, sb));
    HBG_HIP(hipMemcpyAsync(hout.data(), s->d_out.p + nout * slot(h, 0), nb * nout * sizeof(double),
                           hipMemcpyDeviceToHost, sb));
    HBG_HIP(hipStreamSynchronize(sb));
    if (halves == 1 && t + 1 < nbatches) enqueue_batch(t + 1);
    if (prof) {
      s->prof[3] += lus.total_ms;
      s->prof[4] += lus.gemm_ms;
      s->prof[5] += lus.gemm_flops;
      s->prof[6] += (double)lus.gemm_launches;
      s->prof[7] += stage_ms(s->bev[e0], s->bev[e1]);
      s->prof[8] += nb;
      s->prof[9] += nbg * 8.0 / 3.0 * (double)N * N * N;
      s->prof[10] += nbg * (double)N * N * (m32 ? 4 * sizeof(float) + 3 * sizeof(float2)
                                                : sizeof(StaticD) + 2 * sizeof(double2));
      s->prof[12] += (double)lus.fallback_panels;
      s->prof[16] += lus.refine_ms;
      s->prof[17] = std::max(s->prof[17], lus.first_eta);
      s->prof[18] += lus.big_ms;
      s->prof[19] += lus.big_flops;
      s->prof[20] += (double)lus.big_launches;
    }
    for (size_t i = 0; i < nb * nout; ++i)
      if (!std::isfinite(hout[i])) throw Error{HBG_ERR_NUMERICAL, "non-finite solution"};
    for (int b = 0; b < nb; ++b) {
      const int fi = io[f0 + b];
      const double* hA = hout.data() + nout * b;
      const double* hB = hA + nAB;
      const double* hX = hB + nAB;
      const double* hFK = hX + 2 * nX;
      const double* hSC = hFK + 2 * nX;
      const double* hHK = hSC + 2 * nX;
      if (added_mass) std::memcpy(added_mass + fi * nAB, hA, nAB * sizeof(double));
      if (damping) std::memcpy(damping + fi * nAB, hB, nAB * sizeof(double));
      if (excitation) std::memcpy(excitation + fi * 2 * nX, hX, 2 * nX * sizeof(double));
      if (froude_krylov) std::memcpy(froude_krylov + fi * 2 * nX, hFK, 2 * nX * sizeof(double));
      if (scattering) std::memcpy(scattering + fi * 2 * nX, hSC, 2 * nX * sizeof(double));
      if (haskind) std::memcpy(haskind + fi * 2 * nHK, hHK, 2 * nHK * sizeof(double));
      if (s->np > 0)
        std::copy(hfield.begin() + fstride * b, hfield.begin() + fstride * (b + 1), s->field_vals.begin() + fstride * fi);
      if (s->nk > 0)
        std::copy(hkochin.begin() + kstride * b, hkochin.begin() + kstride * (b + 1),
                  s->kochin_vals.begin() + kstride * fi);
    }
  }
  if (s->nk > 0) {
    s->kochin_F = n_omegas;
    s->kochin_H = H;
    s->kochin_nrad = nrad;
  }
  if (s->np > 0) {
    s->field_F = n_omegas;
    s->field_H = H;
    s->field_nrad = nrad;
  }
  // the limits: added mass only (damping and the diffraction quantities are zero there)
  for (int i : il) {
    std::vector<double> a(nAB);
    solve_limit(s, std::isinf(omegas_all[i]), a.data());
    if (added_mass) std::memcpy(added_mass + i * nAB, a.data(), nAB * sizeof(double));
    if (damping) std::fill(damping + i * nAB, damping + (i + 1) * nAB, 0.0);
    if (excitation) std::fill(excitation + i * 2 * nX, excitation + (i + 1) * 2 * nX, 0.0);
    if (froude_krylov) std::fill(froude_krylov + i * 2 * nX, froude_krylov + (i + 1) * 2 * nX, 0.0);
    if (scattering) std::fill(scattering + i * 2 * nX, scattering + (i + 1) * 2 * nX, 0.0);
    if (haskind) std::fill(haskind + i * 2 * nHK, haskind + (i + 1) * 2 * nHK, 0.0);
  }
  if (progress && progress(user, HBG_STAGE_OUTPUTS, n_omegas, n_omegas) == 0)
    throw Error{HBG_ERR_CANCELLED, "cancelled"};
  return HBG_OK;
  GUARD_END("hbg_solve")
}

HbgStatus hbg_solve_group(HbgSession* const* sessions, int32_t count, const double* omegas, int32_t n_omegas,
                          const double* headings_deg, int32_t n_headings, const HbgOptions* options,
                          double* const* added_mass, double* const* damping, double* const* excitation,
                          double* const* froude_krylov, double* const* scattering, double* const* haskind) {
  if (!sessions || count < 1 || (n_omegas > 0 && !omegas) || (n_headings > 0 && !headings_deg))
    return fail(HBG_ERR_ARGUMENT, "null argument or no session", "hbg_solve_group");
  if (n_omegas < 0 || n_headings < 0) return fail(HBG_ERR_ARGUMENT, "negative count", "hbg_solve_group");
  for (int i = 0; i < count; ++i) {
    if (!sessions[i]) return fail(HBG_ERR_ARGUMENT, "null session", "hbg_solve_group");
    if (sessions[i]->N == 0) return fail(HBG_ERR_STATE, "a session has no panels", "hbg_solve_group");
    if (sessions[i]->device != sessions[0]->device)
      return fail(HBG_ERR_VALUE, "the sessions of a group must share their device", "hbg_solve_group");
    for (int j = 0; j < i; ++j)
      if (sessions[j] == sessions[i]) return fail(HBG_ERR_VALUE, "a session twice in the group", "hbg_solve_group");
  }
  for (int i = 0; i < n_omegas; ++i)
    if (!(omegas[i] > 0.0) || !std::isfinite(omegas[i]))
      return fail(HBG_ERR_VALUE, "omega " + std::to_string(i) + " must be finite and > 0", "hbg_solve_group");
  HbgSession* lead = sessions[0];
  std::vector<hipStream_t> own(count);
  for (int i = 0; i < count; ++i) own[i] = sessions[i]->stream;
  // every session's work goes on the lead session's stream for the call (restored on every exit)
  struct StreamRestore {
    HbgSession* const* s;
    std::vector<hipStream_t>& own;
    ~StreamRestore() {
      for (size_t i = 0; i < own.size(); ++i) s[i]->stream = own[i];
    }
  } restore{sessions, own};
  GUARD_BEGIN
  check_options(options);
  HbgOptions o = defaults(options);
  for (int i = 0; i < count; ++i)
    if (o.surrogate && sessions[i]->NI > 0)
      throw Error{HBG_ERR_UNSUPPORTED, "the surrogate with panels in z = 0 (an interior free surface or lids)"};
  use_device(lead);
  const int H = n_headings;
  int nmax = 0, Rmax = 0;
  std::vector<int> Ns(count), nrad(count), Rs(count);
  std::vector<double> heads(std::max(H, 1), 0.0);
  for (int h = 0; h < H; ++h) heads[h] = headings_deg[h] * PI / 180.0;
  for (int i = 0; i < count; ++i) {
    HbgSession* si = sessions[i];
    Ns[i] = si->N;
    nrad[i] = 6 * si->nb;
    Rs[i] = nrad[i] + H;
    if (Rs[i] > 32) throw Error{HBG_ERR_UNSUPPORTED, "more than 32 right-hand sides"};
    nmax = std::max(nmax, Ns[i]);
    Rmax = std::max(Rmax, Rs[i]);
    check_depth(si);
    ensure_statics(si, o.logarithmic ? 1 : 0, o.log_tolerance);
    si->d_heads.ensure(heads.size());
    HBG_HIP(hipMemcpy(si->d_heads.p, heads.data(), heads.size() * sizeof(double), hipMemcpyHostToDevice));
    if (std::isfinite(si->depth) && n_omegas > 0) {  // the frequencies' roots, uploaded once
      std::vector<FinFreq> ffs(n_omegas);
      for (int f = 0; f < n_omegas; ++f) ffs[f] = make_fin_freq(si->depth, omegas[f] * omegas[f] / si->grav);
      si->d_ffs.ensure(ffs.size());
      HBG_HIP(hipMemcpy(si->d_ffs.p, ffs.data(), ffs.size() * sizeof(FinFreq), hipMemcpyHostToDevice));
    }
  }
  for (int i = 0; i < count; ++i) sessions[i]->stream = lead->stream;
  hipStream_t st = lead->stream;
  // the FP32 assembly keeps one geometry's matrices per system; a group runs it as
  // HBG_PRECISION_MIXED (FP64 assembly, complex64 LU, FP64 refinement)
  const bool mixed = o.precision == HBG_PRECISION_MIXED || o.precision == HBG_PRECISION_MIXED32;
  // the units (session, frequency) in decreasing order of N (stable): a batch is padded to its own
  // largest order only, so systems of very different sizes rarely share a batch. A unit is Gs
  // systems in consecutive slots of one batch: the characters of a symmetric body (one with walls
  // or without images); the boundary values and the forces take its NF = Gs N panels.
  std::vector<std::pair<int, int>> order;
  for (int fi = 0; fi < n_omegas; ++fi)
    for (int i = 0; i < count; ++i) order.push_back({i, fi});
  std::stable_sort(order.begin(), order.end(),
                   [&](const std::pair<int, int>& x, const std::pair<int, int>& y) { return Ns[x.first] > Ns[y.first]; });
  const int units = (int)order.size();
  int systems = 0, gmax = 1;
  for (const auto& u : order) {
    systems += sessions[u.first]->Gs;
    gmax = std::max(gmax, sessions[u.first]->Gs);
  }
  const size_t mstride_max = (size_t)nmax * (nmax + Rmax);
  const int B = std::max(batch_size(lead, nmax, Rmax, systems, mixed), gmax);
  lead->d_M.ensure(mstride_max * B);
  if (mixed) {
    lead->d_M32.ensure(mstride_max * B);
    lead->d_D32.ensure((size_t)nmax * Rmax * B);
    lead->d_X64.ensure((size_t)nmax * Rmax * B);
    lead->d_rpart.ensure(residual_max_elems(nmax) * B);
    lead->d_rpart2.ensure(residual_part_elems(nmax, Rmax) * B);
    lead->d_rscale.ensure((size_t)Rmax * B);
    lead->d_active.ensure((size_t)B);
  }
  std::vector<size_t> hstride(count), nout(count);
  for (int i = 0; i < count; ++i) {
    HbgSession* si = sessions[i];
    const int Gs = si->Gs, NF = Gs * Ns[i];
    const size_t D = nrad[i], nAB = D * nrad[i], nX = (size_t)H * D, nHK = (size_t)nrad[i] * H;
    hstride[i] = (size_t)std::max(H, 1) * NF;
    nout[i] = 2 * nAB + 6 * nX + 2 * nHK;
    // S: leading dimension up to nmax, as A (the characters' S_c: leading dimension N)
    si->d_S.ensure(si->G > 1 ? (size_t)Ns[i] * Ns[i] * Gs : (size_t)nmax * Ns[i]);
    si->d_phiI.ensure(hstride[i] * B);
    si->d_dphiI.ensure(hstride[i] * B);
    si->d_u.ensure((size_t)Rs[i] * NF);
    si->d_part.ensure((size_t)rhs_chunks(Ns[i]) * Rs[i] * Ns[i] * Gs);
    si->d_out.ensure(nout[i] * B);
    if (Gs > 1) {
      si->d_uc.ensure((size_t)Ns[i] * Rs[i] * Gs);
      si->d_phif.ensure((size_t)NF * Rs[i]);
    }
  }
  const double logc = o.logarithmic ? 1.0 : 0.0;
  // unit q into slots b .. b + Gs - 1 of a batch of order n (leading dimension n) with R
  // right-hand sides: each system's matrices in its leading N x N block, its right-hand sides at
  // columns n + c, the rest padding
  auto build = [&](int q, int b, int n, int R, size_t mstride) {
    const int i = order[q].first, fi = order[q].second;
    HbgSession* si = sessions[i];
    const int N = Ns[i], Gs = si->Gs, NF = Gs * N;
    const double w = omegas[fi], K = w * w / si->grav;
    const double depth = std::isfinite(si->depth) ? si->depth : 0.0;
    double2* Mb = lead->d_M.p + mstride * b;
    float2* M32b = mixed ? lead->d_M32.p + mstride * b : nullptr;
    const double kw = assemble_frequency(si, K, logc, Mb, o.surrogate != 0, M32b, n, mstride,
                                         std::isfinite(si->depth) ? si->d_ffs.p + fi : nullptr);
    launch_boundary(si->d_panels.p, si->d_body.p, si->d_nj.p, NF, nrad[i], H, w, kw, si->grav, depth, si->d_heads.p,
                    si->d_phiI.p + hstride[i] * b, si->d_dphiI.p + hstride[i] * b, si->d_u.p, st, si->walls);
    if (Gs == 1) {
      launch_rhs(si->d_S.p, N, si->G > 1 ? N : n, si->d_u.p, Rs[i], Mb + (size_t)(n - N) * n, n, si->d_part.p, st);
    } else {
      launch_symmetry_project(si->d_u.p, N, Gs, Rs[i], si->d_chisym.p, si->d_uc.p, st);
      launch_rhs_batched(si->d_S.p, (size_t)N * N, N, N, si->d_uc.p, (size_t)Rs[i] * N, Rs[i],
                         Mb + (size_t)(n - N) * n, mstride, n, si->d_part.p, Gs, st);
    }
    for (int c = 0; c < Gs; ++c)
      launch_pad_system(Mb + mstride * c, M32b ? M32b + mstride * c : nullptr, n, N, R, Rs[i], st);
  };
  std::vector<int> status(B), iters(B), nsys(B), first(B + 1);
  std::vector<double> eta(B), hout;
  std::vector<size_t> hoff(B + 1);
  for (int q0 = 0, nu = 0; q0 < units; q0 += nu) {
    // the batch: the next units (in decreasing order of N) while their systems fit and their
    // order is at least GROUP_ORDER_RATIO of the first one's, so that no system is padded far
    // beyond its own order (the panel and substitution kernels run over the batch's order)
    nu = 0;
    int nb = 0;
    while (q0 + nu < units && nb + sessions[order[q0 + nu].first]->Gs <= B &&
           (nu == 0 || Ns[order[q0 + nu].first] >= GROUP_ORDER_RATIO * Ns[order[q0].first])) {
      first[nu] = nb;
      nb += sessions[order[q0 + nu].first]->Gs;
      ++nu;
    }
    first[nu] = nb;
    int n = 0, R = 0;
    for (int u = 0; u < nu; ++u) {
      n = std::max(n, Ns[order[q0 + u].first]);
      R = std::max(R, Rs[order[q0 + u].first]);
    }
    const size_t mstride = (size_t)n * (n + R);
    for (int u = 0; u < nu; ++u) {
      build(q0 + u, first[u], n, R, mstride);
      for (int b = first[u]; b < first[u + 1]; ++b) nsys[b] = Ns[order[q0 + u].first];
    }
    bool uneven = false;
    for (int b = 0; b < nb; ++b) uneven = uneven || nsys[b] != n;
    const int* ns = uneven ? nsys.data() : nullptr;
    // the unit of a failing system is rebuilt whole (its characters are assembled together), and
    // each of its systems solved again by the plain FP64 path
    auto unit_of = [&](int b) {
      int u = 0;
      while (first[u + 1] <= b) ++u;
      return u;
    };
    if (mixed) {
      mixed_solve_batched(lead->d_M.p, lead->d_M32.p, lead->d_X64.p, lead->d_D32.p, lead->d_rpart.p, lead->d_rpart2.p,
                          lead->d_rscale.p, nb, n, R, true, MIXED_TOL, MIXED_MAX_ITER, st, nullptr, status.data(),
                          iters.data(), eta.data(), ns, lead->d_active.p);
      for (int u = 0; u < nu; ++u) {
        bool fail_u = false;
        for (int b = first[u]; b < first[u + 1]; ++b) fail_u = fail_u || status[b] < 0;
        if (!fail_u) continue;
        build(q0 + u, first[u], n, R, mstride);
        for (int b = first[u]; b < first[u + 1]; ++b)
          lu_solve_batched(lead->d_M.p + mstride * b, mstride, 1, n, n, R, st, nullptr, &status[b],
                           uneven ? &nsys[b] : nullptr);
      }
    } else {
      lu_solve_batched(lead->d_M.p, mstride, nb, n, n, R, st, nullptr, status.data(), ns);
    }
    for (int b = 0; b < nb; ++b) {
      if (status[b] < 0) {  // a panel needs interchanges: rebuild the unit, pivoted path for its systems
        const int u = unit_of(b);
        build(q0 + u, first[u], n, R, mstride);
        for (int c = first[u]; c < first[u + 1]; ++c) {
          const int sc = lu_solve_pivoted(lead->d_M.p + mstride * c, n, n, R, st, nullptr);
          if (c >= b) status[c] = sc;
          else if (sc != 0) status[b] = sc;
        }
      }
      if (status[b] > 0)
        throw Error{HBG_ERR_NUMERICAL, "singular system matrix (zero pivot " + std::to_string(status[b]) + ")"};
    }
    for (int u = 0; u < nu; ++u) {
      const int i = order[q0 + u].first, fi = order[q0 + u].second, b = first[u];
      HbgSession* si = sessions[i];
      const int N = Ns[i], Gs = si->Gs;
      const double2* sol = lead->d_M.p + mstride * b + (size_t)n * n;  // the solution columns, leading dimension n
      if (Gs == 1) {
        launch_forces(si->d_panels.p, si->d_body.p, si->d_nj.p, N, nrad[i], H, omegas[fi], si->rho,
                      lead->d_M.p + mstride * b + (size_t)(n - N) * n, n, si->d_phiI.p + hstride[i] * b,
                      si->d_dphiI.p + hstride[i] * b, si->d_out.p + nout[i] * b, st);
      } else {
        // the potential on the whole body from the characters (lo_frequency.rs:931-943)
        launch_symmetry_recombine(sol, mstride, n, N, Gs, Rs[i], si->d_chisym.p, si->d_phif.p, st);
        launch_forces_sol(si->d_panels.p, si->d_body.p, si->d_nj.p, Gs * N, nrad[i], H, omegas[fi], si->rho,
                          si->d_phif.p, Gs * N, si->d_phiI.p + hstride[i] * b, si->d_dphiI.p + hstride[i] * b,
                          si->d_out.p + nout[i] * b, st);
      }
    }
    // every unit's outputs in one host buffer, one wait for the batch
    hoff[0] = 0;
    for (int u = 0; u < nu; ++u) hoff[u + 1] = hoff[u] + nout[order[q0 + u].first];
    hout.resize(hoff[nu]);
    for (int u = 0; u < nu; ++u) {
      const int i = order[q0 + u].first;
      HBG_HIP(hipMemcpyAsync(hout.data() + hoff[u], sessions[i]->d_out.p + nout[i] * first[u],
                             nout[i] * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    HBG_HIP(hipStreamSynchronize(st));
    for (double v : hout)
      if (!std::isfinite(v)) throw Error{HBG_ERR_NUMERICAL, "non-finite solution"};
    for (int u = 0; u < nu; ++u) {
      const int i = order[q0 + u].first, fi = order[q0 + u].second;
      const size_t D = nrad[i], nAB = D * nrad[i], nX = (size_t)H * D, nHK = (size_t)nrad[i] * H;
      const double* hA = hout.data() + hoff[u];
      const double* hB = hA + nAB;
      const double* hX = hB + nAB;
      const double* hFK = hX + 2 * nX;
      const double* hSC = hFK + 2 * nX;
      const double* hHK = hSC + 2 * nX;
      if (added_mass && added_mass[i]) std::memcpy(added_mass[i] + fi * nAB, hA, nAB * sizeof(double));
      if (damping && damping[i]) std::memcpy(damping[i] + fi * nAB, hB, nAB * sizeof(double));
      if (excitation && excitation[i]) std::memcpy(excitation[i] + fi * 2 * nX, hX, 2 * nX * sizeof(double));
      if (froude_krylov && froude_krylov[i]) std::memcpy(froude_krylov[i] + fi * 2 * nX, hFK, 2 * nX * sizeof(double));
      if (scattering && scattering[i]) std::memcpy(scattering[i] + fi * 2 * nX, hSC, 2 * nX * sizeof(double));
      if (haskind && haskind[i]) std::memcpy(haskind[i] + fi * 2 * nHK, hHK, 2 * nHK * sizeof(double));
    }
  }
  return HBG_OK;
  GUARD_END("hbg_solve_group")
}

HbgStatus hbg_motions(HbgSession* s, int32_t n_omegas, const double* omegas, int32_t n_headings, int32_t n_dof,
                      const double* mass, const double* stiffness, const double* external_damping,
                      const double* added_mass, const double* damping, const double* excitation, double* rao) {
  if (!s || (n_omegas > 0 && (!omegas || !added_mass || !damping || !excitation || !rao)) || !mass || !stiffness)
    return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_motions");
  if (n_omegas < 0 || n_headings < 0 || n_dof < 1)
    return fail(HBG_ERR_VALUE, "negative count or no degree of freedom", "hbg_motions");
  if (n_dof > motions_max_dof())
    return fail(HBG_ERR_UNSUPPORTED, "more than 64 degrees of freedom", "hbg_motions");
  if (n_omegas == 0 || n_headings == 0) return HBG_OK;
  if (n_headings > motions_max_headings()) {
    // the headings in groups of the kernel's right-hand sides (the excitation's rows of each group
    // gathered, its motions scattered back)
    const int hm = motions_max_headings();
    const size_t F = n_omegas, D = n_dof, H = n_headings;
    std::vector<double> x, r;
    for (int h0 = 0; h0 < n_headings; h0 += hm) {
      const size_t hc = std::min<size_t>(hm, H - h0);
      x.assign(2 * F * hc * D, 0.0);
      r.assign(x.size(), 0.0);
      for (size_t f = 0; f < F; ++f)
        std::memcpy(x.data() + 2 * f * hc * D, excitation + 2 * (f * H + h0) * D, 2 * hc * D * sizeof(double));
      const HbgStatus st = hbg_motions(s, n_omegas, omegas, (int32_t)hc, n_dof, mass, stiffness, external_damping,
                                       added_mass, damping, x.data(), r.data());
      if (st != HBG_OK) return st;
      for (size_t f = 0; f < F; ++f)
        std::memcpy(rao + 2 * (f * H + h0) * D, r.data() + 2 * f * hc * D, 2 * hc * D * sizeof(double));
    }
    return HBG_OK;
  }
  {
    // every input finite before anything runs (the elimination would carry a NaN into its pivots)
    const size_t F = n_omegas, D = n_dof, H = n_headings;
    auto finite = [](const double* a, size_t n) {
      for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
      return true;
    };
    if (!finite(omegas, F)) return fail(HBG_ERR_VALUE, "non-finite omega", "hbg_motions");
    if (!finite(mass, D * D)) return fail(HBG_ERR_VALUE, "non-finite mass matrix entry", "hbg_motions");
    if (!finite(stiffness, D * D)) return fail(HBG_ERR_VALUE, "non-finite stiffness matrix entry", "hbg_motions");
    if (external_damping && !finite(external_damping, D * D))
      return fail(HBG_ERR_VALUE, "non-finite external damping entry", "hbg_motions");
    if (!finite(added_mass, F * D * D)) return fail(HBG_ERR_VALUE, "non-finite added mass entry", "hbg_motions");
    if (!finite(damping, F * D * D)) return fail(HBG_ERR_VALUE, "non-finite damping entry", "hbg_motions");
    if (!finite(excitation, 2 * F * H * D)) return fail(HBG_ERR_VALUE, "non-finite excitation entry", "hbg_motions");
  }
  GUARD_BEGIN
  use_device(s);
  const size_t F = n_omegas, D = n_dof, H = n_headings;
  DevBuf<double> dw, dM, dC, dBx, dA, dB;
  DevBuf<double2> dX, dout;
  DevBuf<int> dst;
  dw.ensure(F);
  dM.ensure(D * D);
  dC.ensure(D * D);
  dA.ensure(F * D * D);
  dB.ensure(F * D * D);
  dX.ensure(F * H * D);
  dout.ensure(F * H * D);
  dst.ensure(F);
  HBG_HIP(hipMemcpy(dw.p, omegas, F * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dM.p, mass, D * D * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dC.p, stiffness, D * D * sizeof(double), hipMemcpyHostToDevice));
  if (external_damping) {
    dBx.ensure(D * D);
    HBG_HIP(hipMemcpy(dBx.p, external_damping, D * D * sizeof(double), hipMemcpyHostToDevice));
  }
  HBG_HIP(hipMemcpy(dA.p, added_mass, F * D * D * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dB.p, damping, F * D * D * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dX.p, excitation, F * H * D * sizeof(double2), hipMemcpyHostToDevice));
  launch_motions((int)F, (int)D, (int)H, dw.p, dM.p, dC.p, external_damping ? dBx.p : nullptr, dA.p, dB.p, dX.p,
                 dout.p, dst.p, s->stream);
  std::vector<int> st(F);
  HBG_HIP(hipMemcpyAsync(st.data(), dst.p, F * sizeof(int), hipMemcpyDeviceToHost, s->stream));
  HBG_HIP(hipMemcpyAsync(rao, dout.p, F * H * D * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  for (size_t f = 0; f < F; ++f)
    if (st[f] > 0) {
      char buf[64];
      std::snprintf(buf, sizeof(buf), "%.6g", omegas[f]);
      throw Error{HBG_ERR_NUMERICAL, std::string("motion equation at omega = ") + buf +
                                         ": singular system matrix (zero pivot " + std::to_string(st[f]) + ")"};
    }
  return HBG_OK;
  GUARD_END("hbg_motions")
}

int32_t hbg_surrogate_size(void) { return NN_PARAMS; }

HbgStatus hbg_surrogate_train(HbgSession* s, const HbgSurrogateTraining* t, double* report) {
  if (!s || !t) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_surrogate_train");
  if (t->n_train < t->batch || t->batch < 64 || t->n_holdout < 0 || t->epochs < 1 ||
      !(t->learning_rate > 0.0) || t->n_train > (1 << 28) || t->n_holdout > (1 << 26))
    return fail(HBG_ERR_VALUE, "training sizes out of range", "hbg_surrogate_train");
  GUARD_BEGIN
  use_device(s);
  std::vector<float> P;
  double rep[HBG_SURROGATE_REPORT] = {};
  nn_train(t->seed, t->n_train, t->n_holdout, t->epochs, t->batch, (float)t->learning_rate, s->T, s->ktab, s->fitm,
           s->d_gl.p + 16 * 32, P, rep, s->nn, s->stream);
  if (report) std::memcpy(report, rep, sizeof(rep));
  return HBG_OK;
  GUARD_END("hbg_surrogate_train")
}

HbgStatus hbg_surrogate_set(HbgSession* s, const float* params, int32_t n) {
  if (!s || !params) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_surrogate_set");
  if (n != NN_PARAMS) return fail(HBG_ERR_VALUE, "expected hbg_surrogate_size() parameters", "hbg_surrogate_set");
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(params[i])) return fail(HBG_ERR_VALUE, "non-finite parameter", "hbg_surrogate_set");
  GUARD_BEGIN
  use_device(s);
  s->nn.upload(params);
  return HBG_OK;
  GUARD_END("hbg_surrogate_set")
}

HbgStatus hbg_surrogate_get(HbgSession* s, float* params, int32_t n) {
  if (!s || !params) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_surrogate_get");
  if (!s->nn.ready) return fail(HBG_ERR_STATE, "no surrogate trained or set", "hbg_surrogate_get");
  if (n != NN_PARAMS) return fail(HBG_ERR_VALUE, "expected hbg_surrogate_size() parameters", "hbg_surrogate_get");
  std::memcpy(params, s->nn.params.data(), NN_PARAMS * sizeof(float));
  return HBG_OK;
}

HbgStatus hbg_surrogate_eval(HbgSession* s, int32_t n, const double* in, double* out) {
  if (!s || (n > 0 && (!in || !out))) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_surrogate_eval");
  if (!s->nn.ready) return fail(HBG_ERR_STATE, "no surrogate trained or set", "hbg_surrogate_eval");
  if (n < 0) return fail(HBG_ERR_VALUE, "n must be >= 0", "hbg_surrogate_eval");
  for (int i = 0; i < n; ++i)
    if (!(in[4 * i] > 0.0) || !(in[4 * i + 3] > 0.0) || !std::isfinite(in[4 * i]) || !std::isfinite(in[4 * i + 3]))
      return fail(HBG_ERR_VALUE, "point " + std::to_string(i) + ": R/h and nu h must be finite and > 0",
                  "hbg_surrogate_eval");
  if (n == 0) return HBG_OK;
  GUARD_BEGIN
  use_device(s);
  std::vector<float4> X(n), Y(n);
  for (int i = 0; i < n; ++i) X[i] = nn_input(in[4 * i], in[4 * i + 1], in[4 * i + 2], in[4 * i + 3]);
  DevBuf<float4> dX, dY;
  dX.ensure(n);
  dY.ensure(n);
  HBG_HIP(hipMemcpy(dX.p, X.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  nn_eval(s->nn, dX.p, n, dY.p, s->stream);
  HBG_HIP(hipMemcpyAsync(Y.data(), dY.p, n * sizeof(float4), hipMemcpyDeviceToHost, s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  for (int i = 0; i < n; ++i) {
    double s0, s1;
    nn_scales(in[4 * i], in[4 * i + 3], s0, s1);
    out[4 * i] = Y[i].x * s0;
    out[4 * i + 1] = Y[i].y * s1;
    out[4 * i + 2] = Y[i].z * s1;
    out[4 * i + 3] = Y[i].w * s1;
  }
  return HBG_OK;
  GUARD_END("hbg_surrogate_eval")
}

HbgStatus hbg_finite_remainder(HbgSession* s, double h, double K, int32_t n, const double* R, const double* z,
                               const double* zeta, double* out) {
  if (!s || (n > 0 && (!R || !z || !zeta || !out))) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_finite_remainder");
  if (!(h > 0.0) || !std::isfinite(h) || !(K > 0.0) || !std::isfinite(K) || n < 0)
    return fail(HBG_ERR_VALUE, "h and K must be finite and > 0", "hbg_finite_remainder");
  for (int i = 0; i < n; ++i)
    if (!(R[i] > 0.0) || !(z[i] <= 0.0) || !(zeta[i] <= 0.0) || z[i] < -h || zeta[i] < -h)
      return fail(HBG_ERR_VALUE, "point " + std::to_string(i) + " outside R > 0, -h <= z, zeta <= 0",
                  "hbg_finite_remainder");
  if (n == 0) return HBG_OK;
  GUARD_BEGIN
  use_device(s);
  const FinFreq f = make_fin_freq(h, K);
  std::vector<PanelD> pts(2 * (size_t)n);
  std::memset(pts.data(), 0, pts.size() * sizeof(PanelD));
  for (int i = 0; i < n; ++i) {
    pts[i].c[2] = z[i];
    pts[n + i].c[2] = zeta[i];
  }
  DevBuf<PanelD> dp;
  DevBuf<double> dphi1, dphi2, dR, dz, dzeta, dout;
  DevBuf<double4> dpf;
  DevBuf<double2> dpcs;
  DevBuf<FinFreq> dff;
  dp.ensure(pts.size());
  dphi1.ensure((size_t)FTS * FTW * FTILE);
  dphi2.ensure((size_t)FTS * FTV * FTILE);
  dpf.ensure(pts.size());
  dpcs.ensure(pts.size() * FIN_ROOTS);
  dR.ensure(n);
  dz.ensure(n);
  dzeta.ensure(n);
  dout.ensure(4 * (size_t)n);
  dff.ensure(1);
  HBG_HIP(hipMemcpy(dp.p, pts.data(), pts.size() * sizeof(PanelD), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dR.p, R, n * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dz.p, z, n * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dzeta.p, zeta, n * sizeof(double), hipMemcpyHostToDevice));
  HBG_HIP(hipMemcpy(dff.p, &f, sizeof(FinFreq), hipMemcpyHostToDevice));
  s->d_fwork.ensure(fin_tables_work());
  launch_fin_tables(f, s->T, s->d_gl.p + 16 * 32, dphi1.p, dphi2.p, s->fitm, dp.p, 2 * n, dpf.p, dpcs.p,
                    s->d_fwork.p, s->stream);
  FinTables F;
  F.phi1 = dphi1.p;
  F.phi2 = dphi2.p;
  F.ktab = s->ktab;
  F.pf = dpf.p;
  F.pcs = dpcs.p;
  launch_remainder(dff.p, F, s->T, n, dR.p, dz.p, dzeta.p, dout.p, s->stream);
  HBG_HIP(hipMemcpyAsync(out, dout.p, 4 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HBG_HIP(hipStreamSynchronize(s->stream));
  return HBG_OK;
  GUARD_END("hbg_finite_remainder")
}

HbgStatus hbg_surrogate_gradient(HbgSession* s, const float* params, int32_t batch, const float* inputs,
                                 const float* targets, float* gradient, double* loss) {
  if (!s || !params || !inputs || !targets || !gradient)
    return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_surrogate_gradient");
  if (batch < 1 || batch > (1 << 22)) return fail(HBG_ERR_VALUE, "batch out of range", "hbg_surrogate_gradient");
  GUARD_BEGIN
  use_device(s);
  const double l = nn_gradient(params, inputs, targets, batch, gradient, s->stream);
  if (loss) *loss = l;
  return HBG_OK;
  GUARD_END("hbg_surrogate_gradient")
}

HbgStatus hbg_set_tuning(HbgSession* s, int32_t key, int32_t value) {
  if (!s) return fail(HBG_ERR_ARGUMENT, "session is null", "hbg_set_tuning");
  switch (key) {
    case HBG_TUNE_LU_GEMM:
      if (value < 1 || value > 7) return fail(HBG_ERR_VALUE, "LU GEMM variant must be 1..7", "hbg_set_tuning");
      lu_set_gemm_variant(value);
      return HBG_OK;
    case HBG_TUNE_LU_DEPTH:
      if (value < 1 || value > 64) return fail(HBG_ERR_VALUE, "LU block depth must be in 1..64", "hbg_set_tuning");
      lu_set_depth(value);
      return HBG_OK;
    case HBG_TUNE_LU_GRAPHS:
      if (value != 0 && value != 1) return fail(HBG_ERR_VALUE, "graphs: 0 or 1", "hbg_set_tuning");
      lu_set_graphs(value);
      return HBG_OK;
    case HBG_TUNE_LU_GEMM_MIXED:
      if (value < 1 || value > 7) return fail(HBG_ERR_VALUE, "LU GEMM variant must be 1..7", "hbg_set_tuning");
      lu_set_gemm_variant32(value);
      return HBG_OK;
    case HBG_TUNE_LU_DEPTH_MIXED:
      if (value < 1 || value > 64) return fail(HBG_ERR_VALUE, "LU block depth must be in 1..64", "hbg_set_tuning");
      lu_set_depth32(value);
      return HBG_OK;
    case HBG_TUNE_MIXED32_TOL:
      if (value < 6 || value > 14) return fail(HBG_ERR_VALUE, "refinement tolerance exponent must be in 6..14",
                                               "hbg_set_tuning");
      g_mixed32_tol_exp = value;
      return HBG_OK;
    case HBG_TUNE_LU_SUB:
    case HBG_TUNE_LU_SUB_MIXED:
    case HBG_TUNE_LU_SUB2:
    case HBG_TUNE_LU_SUB2_MIXED:
      if (value < 0 || value > 64) return fail(HBG_ERR_VALUE, "LU sub-block: 0..64 panels", "hbg_set_tuning");
      if (key == HBG_TUNE_LU_SUB)
        lu_set_sub(value);
      else if (key == HBG_TUNE_LU_SUB_MIXED)
        lu_set_sub32(value);
      else if (key == HBG_TUNE_LU_SUB2)
        lu_set_sub2(value);
      else
        lu_set_sub2_32(value);
      return HBG_OK;
    case HBG_TUNE_NN_RMIN:
      if (value < 10 || value > 2000)
        return fail(HBG_ERR_VALUE, "the surrogate's lower R/h bound must be in 10..2000 hundredths", "hbg_set_tuning");
      s->nn_rmin = value / 100.0;
      return HBG_OK;
    default:
      return fail(HBG_ERR_VALUE, "unknown tuning key " + std::to_string(key), "hbg_set_tuning");
  }
}

HbgStatus hbg_get_tuning(HbgSession* s, int32_t key, int32_t* value) {
  if (!s || !value) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_get_tuning");
  switch (key) {
    case HBG_TUNE_LU_GEMM:
      *value = lu_gemm_variant();
      return HBG_OK;
    case HBG_TUNE_LU_DEPTH:
      *value = lu_depth();
      return HBG_OK;
    case HBG_TUNE_LU_GRAPHS:
      *value = lu_graphs();
      return HBG_OK;
    case HBG_TUNE_LU_GEMM_MIXED:
      *value = lu_gemm_variant32();
      return HBG_OK;
    case HBG_TUNE_LU_DEPTH_MIXED:
      *value = lu_depth32();
      return HBG_OK;
    case HBG_TUNE_MIXED32_TOL:
      *value = g_mixed32_tol_exp;
      return HBG_OK;
    case HBG_TUNE_NN_RMIN:
      *value = (int32_t)std::lround(s->nn_rmin * 100.0);
      return HBG_OK;
    case HBG_TUNE_LU_SUB:
      *value = lu_sub();
      return HBG_OK;
    case HBG_TUNE_LU_SUB_MIXED:
      *value = lu_sub32();
      return HBG_OK;
    case HBG_TUNE_LU_SUB2:
      *value = lu_sub2();
      return HBG_OK;
    case HBG_TUNE_LU_SUB2_MIXED:
      *value = lu_sub2_32();
      return HBG_OK;
    default:
      return fail(HBG_ERR_VALUE, "unknown tuning key " + std::to_string(key), "hbg_get_tuning");
  }
}

HbgStatus hbg_autotune(HbgSession* s, int32_t n, int32_t nrhs, int32_t precision, double* report, int32_t max_rows,
                       int32_t* rows) {
  if (!s || (report && !rows)) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_autotune");
  if (n < 64 || nrhs < 0 || nrhs > 32 || (precision != HBG_PRECISION_FP64 && precision != HBG_PRECISION_MIXED))
    return fail(HBG_ERR_VALUE, "n >= 64, 0 <= nrhs <= 32 and a known precision", "hbg_autotune");
  GUARD_BEGIN
  use_device(s);
  size_t free_b = 0, total_b = 0;
  HBG_HIP(hipMemGetInfo(&free_b, &total_b));
  int r = 0;
  const int b = lu_autotune(precision == HBG_PRECISION_FP64 ? 1 : 0, n, nrhs, 48, free_b / 3, s->stream, report,
                            report ? max_rows : 0, &r);
  if (rows) *rows = r;
  s->max_batch = b;
  return HBG_OK;
  GUARD_END("hbg_autotune")
}

HbgStatus hbg_set_batch(HbgSession* s, int32_t max_batch) {
  if (!s) return fail(HBG_ERR_ARGUMENT, "session is null", "hbg_set_batch");
  if (max_batch < 0 || max_batch > 64) return fail(HBG_ERR_VALUE, "max_batch must be in 0..64", "hbg_set_batch");
  s->max_batch = max_batch;
  return HBG_OK;
}

HbgStatus hbg_set_profiling(HbgSession* s, int32_t enable) {
  if (!s) return fail(HBG_ERR_ARGUMENT, "session is null", "hbg_set_profiling");
  s->profiling = enable != 0;
  for (double& v : s->prof) v = 0.0;
  return HBG_OK;
}

HbgStatus hbg_profile_read(HbgSession* s, double* out, int32_t n) {
  if (!s || !out) return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_profile_read");
  for (int i = 0; i < n && i < HBG_PROFILE_COUNTERS; ++i) out[i] = s->prof[i];
  for (int i = HBG_PROFILE_COUNTERS; i < n; ++i) out[i] = 0.0;
  return HBG_OK;
}

HbgStatus hbg_write_wamit(const char* stem, int32_t overwrite, double density, double gravity,
                          double reference_length, const double* omegas, int32_t n_omegas,
                          const double* headings_deg, int32_t n_headings, int32_t n_bodies,
                          const double* added_mass, const double* damping, const double* haskind,
                          const double* excitation, const char* stamp) {
  if (!stem || !omegas || !added_mass || !damping || (n_headings > 0 && (!headings_deg || !haskind || !excitation)))
    return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_write_wamit");
  GUARD_BEGIN
  std::string err;
  int st = write_wamit(stem, overwrite, density, gravity, reference_length, omegas, n_omegas, headings_deg,
                       n_headings, n_bodies, added_mass, damping, haskind, excitation, nullptr, nullptr, nullptr,
                       stamp, err);
  if (st) throw Error{st, err};
  return HBG_OK;
  GUARD_END("hbg_write_wamit")
}

HbgStatus hbg_write_wamit_files(const char* stem, int32_t overwrite, double density, double gravity,
                                double reference_length, const double* omegas, int32_t n_omegas,
                                const double* headings_deg, int32_t n_headings, int32_t n_bodies,
                                const double* added_mass, const double* damping, const double* haskind,
                                const double* excitation, const double* froude_krylov, const double* scattering,
                                const double* rao, const char* stamp) {
  if (!stem || !omegas || !added_mass || !damping || (n_headings > 0 && (!headings_deg || !haskind || !excitation)))
    return fail(HBG_ERR_ARGUMENT, "null argument", "hbg_write_wamit_files");
  GUARD_BEGIN
  std::string err;
  int st = write_wamit(stem, overwrite, density, gravity, reference_length, omegas, n_omegas, headings_deg,
                       n_headings, n_bodies, added_mass, damping, haskind, excitation, froude_krylov, scattering, rao,
                       stamp, err);
  if (st) throw Error{st, err};
  return HBG_OK;
  GUARD_END("hbg_write_wamit_files")
}

HbgStatus hbg_write_wamit_all(const char* stem, int32_t overwrite, double density, double gravity,
                              double reference_length, const double* omegas, int32_t n_omegas,
                              const double* headings_deg, int32_t n_headings, int32_t n_bodies,
                              const double* added_mass, const double* damping, const double* haskind,
                              const double* excitation, const double* froude_krylov, const double* scattering,
                              const double* rao, int32_t n_points, const double* points,
                              const double* field_radiation, const double* field_diffraction, const char* stamp) {
  if (!stem || !omegas || !added_mass || !damping || (n_headings > 0 && (!headings_deg || !haskind || !excitation)) ||
      n_points < 0 || (n_points > 0 && !points))
    return fail(HBG_ERR_ARGUMENT, "null argument or negative point count", "hbg_write_wamit_all");
  GUARD_BEGIN
  std::string err;
  int st = write_wamit(stem, overwrite, density, gravity, reference_length, omegas, n_omegas, headings_deg,
                       n_headings, n_bodies, added_mass, damping, haskind, excitation, froude_krylov, scattering, rao,
                       stamp, err, n_points, points, field_radiation, n_headings > 0 ? field_diffraction : nullptr);
  if (st) throw Error{st, err};
  return HBG_OK;
  GUARD_END("hbg_write_wamit_all")
}

}  // extern "C"
