/* FP64 CPU oracle of the low-order frequency-domain solve: test infrastructure only.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it (as the checker).
 * Status codes follow the reference's HbStatus values (crates/hydrobem-c/include/hydrobem.h:42-98). */
#ifndef HB_ORACLE_H
#define HB_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif
enum { ORC_OK = 0, ORC_ERR_VALUE = 1, ORC_ERR_UNSUPPORTED = 3, ORC_ERR_ARGUMENT = 7, ORC_ERR_NUMERICAL = 8 };
int orc_wave_kernel(double X, double Y, double* out7);
double orc_dispersion_x0(double H);
double orc_dispersion_xn(double H, int n);
int orc_finite_kernel(double h, double K, double R, double z, double zeta, double* out8);
int orc_finite_phi(double H, int which, double xb, double vb, double* out3);
int orc_finite_limit_kernel(int mode, double h, double R, double z, double zeta, double* out8);
int orc_v_vx(double X, double Y, double* out2);
int orc_yt(double X, double* out2);
int orc_struve01(double x, double* out2);
int orc_gauss_legendre(int n, double* x, double* w);
int orc_make_panel(const double* v12, double* out20);
int orc_rankine(const double* v12, const double* x, double* out2);
int orc_log_singularity(const double* v12, const double* image, double* out2);
void* orc_problem_create(const double* vertices, int n, const int* body, int nbodies, const double* rotation_centres,
                         double rho, double grav, double depth, int logint, double logtol, int* status);
void orc_problem_free(void* p);
int orc_statics(void* p, double* out);
int orc_assemble(void* p, double omega, double* S, double* A);
int orc_lu_solve(int n, double* a, int nrhs, double* b);
int orc_solve(void* p, const double* omegas, int nf, const double* headings_deg, int nh, double* A, double* B,
              double* X, double* FK, double* SC, double* HK);
#ifdef __cplusplus
}
#endif
#endif
