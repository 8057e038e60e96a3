/* pgm_oracle: test infrastructure only.
 *
 * A clean-room FP64 CPU restatement of the reference engine core/pgm_core.cpp (pGM electrostatics:
 * Gaussian charges, covalent permanent dipoles, Gaussian induced dipoles, smooth PME of charges and
 * dipoles, Jacobi-PCG induction), used as the checker of the HIP engine.  Only tests/, the smoke check and
 * the bench's CPU-baseline leg load it; the product path never does.
 *
 * The same C ABI is exported by the recipe-built reference library oracle/_ref/libpgm_ref.so (a driver
 * over the reference's own core/pgm_core.cpp), so a test can run either behind one wrapper.
 *
 * Units: nm, e, e nm, nm^3, kJ/mol.  Box: rows a, b, c in reduced triclinic form (a = (ax,0,0),
 * b = (bx,by,0)), row-major 3x3.  Every function returns 0 on success, -1 on error (orc_last_error()).
 */
#ifndef PGM_ORACLE_H_
#define PGM_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors pgmcore::Settings (core/pgm_core.h:22-34). */
typedef struct {
    int periodic;      /* PME (1) or all pairs without a box (0) */
    double cutoff;     /* direct-space cutoff [nm] */
    double beta;       /* Ewald coefficient [1/nm] */
    int grid[3];       /* PME grid */
    int order;         /* B-spline order */
    int polarizable;   /* induced dipoles on (1) / off (0) */
    double tol;        /* max|alpha r| / mean|alpha E_perm| */
    int maxIter;
    double peek;       /* final mu += peek alpha r */
    int predictor;     /* 0 none, 1 mu3, 2 mu4 */
    double ke;         /* Coulomb constant [kJ mol^-1 nm e^-2] */
} orc_settings;

const char* orc_last_error(void);
const char* orc_kind(void);   /* "port" (this restatement) or "reference" (oracle/_ref) */

/* Engine (core/pgm_core.h:47-115).  Topology arrays are copied. */
void* orc_create(const orc_settings* s, int n, const double* q, const double* radius, const double* alpha,
                 int ncov, const int* covAtom, const int* covPartner, const double* covStrength);
void orc_destroy(void* h);
/* Energy [kJ/mol]; forces (3n, may be NULL) are added to. */
int orc_compute(void* h, const double* pos, const double* box9, double* forcesAdd, double* energy);
int orc_terms(void* h, double* out5);   /* direct, reciprocal, self, background, polarization */
int orc_induced(void* h, double* mu);   /* 3n */
int orc_permanent(void* h, double* p);  /* 3n */
int orc_iterations(void* h);
double orc_residual(void* h);
int orc_strain(void* h, double* out9);
int orc_set_history_key(void* h, long long key);
int orc_reset_history(void* h);
/* Push one solution (3n, most recent last pushed) into the predictor history, as a finished step would. */
int orc_push_history(void* h, const double* mu);
int orc_set_topology(void* h, const double* q, const double* radius, const double* alpha, int ncov,
                     const int* covAtom, const int* covPartner, const double* covStrength);
int orc_set_geometry(void* h, const double* pos, const double* box9);
/* Field [e/nm^2] of charges q (n) and dipoles d (3n) at every atom, after orc_set_geometry. */
int orc_field(void* h, const double* q, const double* d, double* E);
long long orc_num_pairs(void* h);
/* CG preconditioner (options of the reference's CUDA engine, core/cuda/pgm_cuda.cu:1526-1588,1753-1760): the
 * short-range M = D + D T_s D over the pairs within radius (0: Jacobi; off when an atom has more than 48 such
 * pairs), Jacobi with the self term alpha/(1 - alpha c_self).  orc_precondition applies M (after set_geometry);
 * orc_preconditioner_on: 1 when the short-range M was used (the last solve or orc_precondition). */
int orc_set_preconditioner(void* h, double radius, int jacobi_self);
int orc_precondition(void* h, const double* r, double* z);
int orc_preconditioner_on(void* h);

/* Numerics (core/pgm_math.h). */
int orc_erf_kernels(double a, double r, int nmax, double* B);
int orc_erfc_kernels(double a, double r, int nmax, double* B);
int orc_pair_kernels(double a, double beta, double r, int nmax, double* G);
int orc_bspline_moduli(int K, int p, double* out);
double orc_ewald_coefficient(double rc, double rtol);
/* Unnormalized forward r2c (K1 x K2 x (K3/2+1) complex, interleaved re/im) and backward c2r. */
/* GVDW, the Gaussian-density van der Waals (core/gvdw_kernel.h): G(y) = F(y)/y^6 and H(y) = G'(y)/y; the pair
 * energy with dEdR = -(1/r) dU/dr (bij = 1/sqrt(2 (R_i^2 + R_j^2)), rep 0 Gaussian / 1 Slater). */
int orc_gvdw_g(double y, double* G, double* H);
int orc_gvdw_pair(double r, double bij, double A, double C6, double B, int rep, double* energy, double* dEdR);
/* GVDW by the combination rule (A_ij = a_i a_j, C6_ij = c_i c_j, B_ij = (b_i + b_j)/2) over all minimum-image
 * pairs within rvdw except the skipped pairs (skipI[k], skipJ[k]), brute force: energy [kJ/mol], forces added to
 * forcesAdd (3n), strain S_ab = dE/d eps_ab added to strain9 (each may be NULL). */
int orc_gvdw_sum(int n, const double* pos, const double* box9, const double* radius, const double* sqrtA,
                 const double* sqrtC6, const double* b, int rep, double rvdw, int nskip, const int* skipI,
                 const int* skipJ, double* energy, double* forcesAdd, double* strain9);
int orc_fft3d_r2c(int K1, int K2, int K3, const double* in, double* out);
int orc_fft3d_c2r(int K1, int K2, int K3, const double* in, double* out);

#ifdef __cplusplus
}
#endif

#endif
