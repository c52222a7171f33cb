/*
 * ehd_oracle.h -- FP64 CPU oracle for the coupled EHD time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * It restates, in plain C with runtime grid sizes, the reference's
 * coupled step (main.cu:181-208):
 *   stream_collide_save  LBM.cu:278-293   (collide, boundary, stream, bc_charge)
 *   extension            poisson.cu:78-121
 *   3-D Z2Z DFT          main.cu:194,200  (own FFT; cuFFT's unnormalised convention)
 *   derivative           poisson.cu:123-143
 *   extract              poisson.cu:145-167
 *   efield + gpu_bc      poisson.cu:32-73
 * with the reference's data layout: scalar fields [z][y][x], moving
 * populations [d-1][z][y][x] (LBM.cu:21-34).
 */
#ifndef EHD_ORACLE_H
#define EHD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Step parameters (LBM.h:31-80 plus the runtime grid). */
typedef struct oref_params {
    int nx, ny, nz, ne;              /* ne = 2*(nz-1)  (LBM.h:37) */
    double lx, ly, lz, dx, dy, dz;   /* LBM.h:39-45; dx=dy=dz=lz/(nz-1), lx=nx*dx */
    double cfl, dt, cs_square;       /* LBM.h:51-53 */
    double rho0, charge0, voltage, eps, diffu, nu, K;   /* LBM.h:54-62 */
    double uw, exf;                  /* LBM.h:47-50 (cross-flow drivers) */
    double V, VC;                    /* TRT magic numbers, flow/charge (LBM.h:79-80) */
    double w0, ws, wa, wd;           /* LBM.h:73-76 */
} oref_params;

/* Derive the dimensional parameters from the nondimensional groups with the
 * inverse of compute_parameters (LBM.cu:1356-1379):
 *   K = sqrt(eps/rho0)/M, nu = eps*V/(K*rho0*T), q0 = C*V*eps/Lz^2, D = K*V/Fe. */
int oref_params_init(oref_params* p, int nx, int ny, int nz,
                     double T, double M, double C, double Fe,
                     double lambda_f, double lambda_c,
                     double cfl, double voltage, double eps, double rho0, double lz,
                     double uw, double exf);

/* Forward map of compute_parameters (LBM.cu:1375-1378). */
void oref_compute_parameters(const oref_params* p, double* T, double* M, double* C, double* Fe);

typedef struct oref_box oref_box;

oref_box* oref_create(const oref_params* p);
void      oref_destroy(oref_box* b);
const oref_params* oref_get_params(const oref_box* b);

/* Field accessors: which = one of OREF_F_* below.  Sizes: scalars S=nx*ny*nz,
 * moving populations 26*S, f0bc 2*nx*ny, spectral buffers 2*nx*ny*ne doubles. */
enum {
    OREF_F_F0 = 0, OREF_F_F1, OREF_F_F2, OREF_F_H0, OREF_F_H1, OREF_F_H2,
    OREF_F_RHO, OREF_F_Q, OREF_F_PHI, OREF_F_UX, OREF_F_UY, OREF_F_UZ,
    OREF_F_EX, OREF_F_EY, OREF_F_EZ, OREF_F_F0BC,
    OREF_F_SEXT, OREF_F_FREQ, OREF_F_PHIEXT, OREF_F_KX, OREF_F_KY, OREF_F_KZ,
    OREF_F_COUNT
};
double* oref_field(oref_box* b, int which);

/* Initial states. */
void oref_initialization(oref_box* b);                 /* LBM.cu:80-96 */
void oref_set_fields(oref_box* b, const double* f9);   /* read_data semantics (LBM.cu:1596-1643): rho,q,phi,u,v,w,Ex,Ey,Ez */
/* Hydrostatic charge and potential plus a seeded perturbation of u in the
 * interior; hydro_density = 1 also sets the hydrostatic pressure (rho0 +
 * (p - mean p)/cs^2, dp/dz = q Ez), 0 keeps rho = rho0 (gpu_initialization). */
void oref_init_hydrostatic(oref_box* b, uint64_t seed, double amp, int hydro_density);
/* as above with the velocity perturbation pattern (0 noise, 1 rolls, 2 squares,
 * 3 crossed rolls, 4 hexagons) */
void oref_init_hydrostatic_pattern(oref_box* b, uint64_t seed, double amp, int hydro_density, int pattern);
void oref_pattern_velocity(int pattern, double amp, double xx, double yy, double zz, double lx, double ly,
                           double lz, double* u);
void oref_init_equilibrium(oref_box* b);               /* LBM.cu:98-276 */

/* LBM phases (LBM.cu:278-1353). */
void oref_collide_save(oref_box* b);   /* LBM.cu:295-1091 */
void oref_boundary(oref_box* b);       /* LBM.cu:1093-1203 */
void oref_stream(oref_box* b);         /* LBM.cu:1205-1279 */
void oref_bc_charge(oref_box* b);      /* LBM.cu:1281-1353 */
void oref_stream_collide_save(oref_box* b);  /* LBM.cu:278-293 */

/* Poisson phases (poisson.cu, main.cu:191-206). */
void oref_extension(oref_box* b);      /* poisson.cu:87-121 */
void oref_fft3d(oref_box* b, double* in, double* out, int sign);  /* cufftExecZ2Z, unnormalised */
void oref_derivative(oref_box* b);     /* poisson.cu:132-143 (acts on FREQ) */
void oref_extract(oref_box* b);        /* poisson.cu:154-167 */
void oref_efield(oref_box* b);         /* poisson.cu:32-73 */
void oref_poisson(oref_box* b);        /* extension -> FFT -> derivative -> IFFT -> extract */
/* Same solution by an x,y DFT plus a per-mode tridiagonal solve in z (the
 * identity the HIP path relies on); used to pin that identity. */
void oref_poisson_tridiag(oref_box* b);

/* One coupled step = main.cu:184-208; n steps. */
void oref_step(oref_box* b, int nsteps);
/* The same step with oref_poisson_tridiag as the Poisson solve: two FP64
 * evaluations of one discrete problem, used to measure the FP64 floor of the
 * coupled dynamics (how far rounding-level differences in phi carry). */
void oref_step_tridiag(oref_box* b, int nsteps);

/* Collector current (LBM.cu:1646-1671) on copies of q and Ez (box unchanged). */
double oref_current(oref_box* b);

/* 1-D unnormalised DFT on interleaved complex data (sign -1 forward, +1 inverse). */
void oref_fft1d(double* data, int n, int sign);

/* Hash used for the seeded perturbation: U(-1,1) from (seed, x, y, z, comp). */
double oref_hash_uniform(uint64_t seed, uint32_t x, uint32_t y, uint32_t z, uint32_t comp);

/* Hydrostatic profile (D=0, u=0) of unipolar injection: nondimensional E0 for
 * injection number C, and q, phi, Ez at height zeta in [0,1]. */
double oref_hydrostatic_e0(double C);

/* Lattice tables of this oracle, exported so tests can compare them with the
 * tables parsed from the reference source text:
 *   c[27*3]      direction vectors
 *   opp[27]      opposite direction
 *   w[27]        weights
 *   eqf[27*4], eqh[27*4]  signed axis terms (+-1..3, 0 = none) of c.u in the
 *                equilibria, in the reference's evaluation order
 *   mom[3*18]    per axis: 9 directions with c=+1 then 9 with c=-1, in summation order */
void oref_tables(int* c, int* opp, double* w, int* eqf, int* eqh, int* mom);

/* Single-node terms of the collision, exactly as oref_collide_save uses them:
 * flow and charge equilibria (fe[27], he[27], LBM.cu:519-660) and the Guo
 * force populations (fpop[27], LBM.cu:664-702). */
void oref_node_terms(const oref_params* p, double rho, double ux, double uy, double uz,
                     double charge, double Ex, double Ey, double Ez,
                     double Fx, double Fy, double Fz, double* fe, double* he, double* fpop);

/* TRT rates of gpu_collide_save (LBM.cu:298-302): omega_plus, omega_minus,
 * omega_c_plus, omega_c_minus. */
void oref_rates(const oref_box* b, double* out4);

/* Override the four rates (same order).  Tests use it to plant a mistake in
 * the oracle on the CPU -- e.g. swapping omega_c_plus and omega_c_minus is
 * the same as swapping their uses in every relaxation line -- and to show
 * that the moment checks fail on it. */
void oref_set_rates(oref_box* b, const double* in4);

/* Set OpenMP thread count used by the oracle (0 = leave default). */
void oref_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
