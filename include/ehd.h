/*
 * ehd.h -- C-ABI of the MI355X implementation of the coupled EHD time step
 * (D3Q27 TRT lattice Boltzmann for flow f and charge h, FFT Poisson solve for
 * the potential, E = -grad(phi)).
 *
 * The reference has no FFI layer: its drop-in boundary is the set of host
 * wrappers main() calls.  Each entry point below names the reference
 * interface it replaces.  All functions return 0 on success and a negative
 * EHD_E* code on failure; ehd_last_error() then holds a message (thread
 * local).  Calls on one box must be serialised by the caller; distinct boxes
 * may be driven from distinct threads.  Host buffers are caller owned, in the
 * reference layout: scalar fields [z][y][x] (x fastest), moving populations
 * [d-1][z][y][x] (LBM.cu:21-34), all double precision.
 */
#ifndef EHD_H
#define EHD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EHD_ABI_VERSION 2

/* error codes */
#define EHD_OK             0
#define EHD_EINVAL        -1   /* bad argument or parameter */
#define EHD_EHIP          -2   /* HIP runtime error */
#define EHD_ENOMEM        -3   /* device allocation failed */
#define EHD_ESTATE        -4   /* call not valid in the box's state */
#define EHD_ENONFINITE    -5   /* NaN/Inf detected by the diagnostics watchdog */

/* storage / arithmetic of the populations */
#define EHD_FP64           0   /* double storage, double arithmetic, reference operation order */
#define EHD_FP32S          1   /* float storage of f - w*rs(z), h - w*qs(z) (hydrostatic shifts), float arithmetic */

/* initial state */
#define EHD_INIT_REFERENCE   0 /* rho0, q=0, linear phi, u=0, E=0 (gpu_initialization, LBM.cu:80-96) */
#define EHD_INIT_HYDROSTATIC 1 /* analytic unipolar-injection profile q(z), phi(z), E=-grad(phi),
                                  plus seeded uniform perturbation of u in the interior */
#define EHD_INIT_FIELDS      2 /* caller-supplied rho,q,phi,u,v,w,Ex,Ey,Ez (read_data, LBM.cu:1596-1643) */

/* perturbation of the velocity in EHD_INIT_HYDROSTATIC (amplitude perturb_amp);
 * zeta = z/Lz, xi = x/Lx, eta = y/Ly */
#define EHD_PERTURB_NOISE    0 /* seeded uniform noise U(-1,1) per node and component */
#define EHD_PERTURB_ROLLS    1 /* x rolls: uz = (cos 2pi zeta - 1) cos 2pi xi, ux = Lx/Lz sin 2pi zeta sin 2pi xi
                                  (the roll pattern of LBM.cu:391-394) */
#define EHD_PERTURB_SQUARES  2 /* square cells (LBM.cu:396-399) */
#define EHD_PERTURB_CROSS    3 /* superposed x and y rolls, the pattern LBM.cu:400-402 applies */
#define EHD_PERTURB_HEXAGONS 4 /* hexagonal cells (LBM.cu:404-411), lateral scale L = Lx/sqrt(3) */

/* Driver surface of main.cu (grid, T, C, M, Fe, Pr, TRT magic parameter) plus
 * the physical constants LBM.h fixes at compile time (LBM.h:31-80). */
typedef struct ehd_params {
    int    nx, ny, nz;        /* grid nodes; walls at z=0 and z=nz-1; periodic x,y (LBM.h:34-36) */
    double T, C, M, Fe;       /* nondimensional groups (compute_parameters, LBM.cu:1356-1379) */
    double Pr;                /* carried as metadata; T, C, M, Fe fix nu and D */
    double lambda_f;          /* TRT magic parameter, flow   (LBM.h:79, 1/12 in the reference) */
    double lambda_c;          /* TRT magic parameter, charge (LBM.h:80, 1e-6) */
    double cfl;               /* dt/dx (LBM.h:51, 0.01) */
    double voltage, eps, rho0, lz;   /* LBM.h:42,54-57 (1e4, 1e-4, 1600, 1) */
    double uw, exf;           /* moving-top-wall speed, x body force (LBM.h:47-50; 0 as shipped) */
    int    precision;         /* EHD_FP64 | EHD_FP32S */
    int    init_mode;         /* EHD_INIT_* */
    uint64_t seed;            /* perturbation seed (EHD_INIT_HYDROSTATIC) */
    double perturb_amp;       /* perturbation amplitude of u, physical units (1e-3) */
    int    hydro_density;     /* EHD_INIT_HYDROSTATIC: 0 = uniform rho0 (as gpu_initialization),
                                 1 = hydrostatic pressure, rho0 + (p - mean p)/cs^2 with dp/dz = q Ez */
    int    perturb_pattern;   /* EHD_PERTURB_* (EHD_INIT_HYDROSTATIC) */
} ehd_params;

/* Dimensional constants derived from ehd_params (inverse of compute_parameters). */
typedef struct ehd_derived {
    double dx, dy, dz, dt, lx, ly, cs_square;
    double K, nu, diffu, charge0;
    double tau_plus, tau_minus, tau_c_plus, tau_c_minus;
} ehd_derived;

/* Global FP64 diagnostics, reduced on the device. */
typedef struct ehd_diag {
    double current;       /* collector current, current() (LBM.cu:1646-1671) */
    double total_charge;  /* sum of q over interior nodes times dx*dy*dz */
    double umax;          /* max of uz with the top-wall extrapolation (record_umax, LBM.cu:1673-1713) */
    double nusselt;       /* current / hydrostatic current of the same box (0 if not set) */
    double mean_rho;      /* mean density over interior nodes */
    long long step;       /* steps taken since init */
    double time;          /* physical time t (main.cu:208) */
    long long nonfinite;  /* nodes whose q, rho, ux, uy or uz is NaN or infinite (the watchdog;
                             ehd_diagnostics then returns EHD_ENONFINITE, with every field filled) */
} ehd_diag;

typedef struct ehd_box ehd_box;

/* Reference defaults (LBM.h): 122x122x101, T=996, M=10, C=10, Fe=4000,
 * lambda 1/12 and 1e-6, FP64, reference initial state. */
int ehd_params_default(ehd_params* p);

/* Inverse of compute_parameters (LBM.cu:1356-1379) plus the TRT rates of
 * gpu_collide_save (LBM.cu:298-302).  Host only, needs no GPU. */
int ehd_derive(const ehd_params* p, ehd_derived* d);

/* Forward map, compute_parameters (LBM.cu:1356-1379). */
int ehd_compute_parameters(const ehd_derived* d, const ehd_params* p,
                           double* T, double* M, double* C, double* Fe);

/* Allocates every device buffer of one box (main.cu:79-107 allocations; the
 * cuFFT plan of main.cu:104 becomes the library's own FFT plans). */
int ehd_create(const ehd_params* p, ehd_box** out);
void ehd_destroy(ehd_box* b);   /* main.cu:283-306 frees */

/* Initial macroscopic state plus equilibrium populations:
 * initialization/read_data then init_equilibrium (main.cu:153-164,
 * LBM.cu:69-276).  fields9 (9*S doubles: rho,q,phi,u,v,w,Ex,Ey,Ez) is
 * required for EHD_INIT_FIELDS and ignored otherwise. */
int ehd_init(ehd_box* b, const double* fields9);

/* Consistent start from given fields (no reference counterpart; after Mei et
 * al., J. Comput. Phys. 219 (2006) 2011, for warm starts): ehd_init with
 * EHD_INIT_FIELDS, then `iterations` coupled steps in which every interior
 * node collides towards the equilibria of the given velocity and charge
 * (its density free, the Poisson solve on the given charge), so that the
 * populations gain the non-equilibrium parts of that state; the equilibrium
 * start alone lacks them and its relaxation excites the slow modes.  The
 * iterations count as steps (step and t advance).  FP32S boxes with
 * init_mode EHD_INIT_FIELDS. */
int ehd_init_consistent(ehd_box* b, const double* fields9, int iterations);

/* nsteps coupled steps, each = stream_collide_save + extension + FFT +
 * derivative + IFFT + extract + efield (main.cu:181-208). */
int ehd_step(ehd_box* b, int nsteps);

/* The LBM part alone, stream_collide_save (LBM.cu:278-293), with the E field
 * of the current potential. */
int ehd_lbm_step(ehd_box* b);

/* The Poisson part alone on the box's current charge: extension, FFT,
 * derivative, IFFT, extract (poisson.cu:78-167, main.cu:191-203). */
int ehd_poisson(ehd_box* b);

/* Copy macroscopic fields to the host; any pointer may be NULL.  Values are
 * those of the last step (the wall rows are the raw collide outputs, as the
 * reference stores them). E is -grad(phi) of the current potential
 * (efield, poisson.cu:32-73). */
int ehd_read_fields(ehd_box* b, double* rho, double* q, double* phi,
                    double* ux, double* uy, double* uz,
                    double* ex, double* ey, double* ez);

/* Overwrite the potential (S doubles) or the charge (S doubles). */
int ehd_write_phi(ehd_box* b, const double* phi);
int ehd_write_charge(ehd_box* b, const double* q);

/* Pre-collision populations between steps, i.e. the reference's f0, f1, h0,
 * h1 after stream and bc_charge (LBM.cu:278-293).  f0,h0: S doubles; f1,h1:
 * 26*S doubles.  Any pointer may be NULL. */
int ehd_read_populations(ehd_box* b, double* f0, double* f1, double* h0, double* h1);
int ehd_write_populations(ehd_box* b, const double* f0, const double* f1,
                          const double* h0, const double* h1);

/* Lossless checkpoint of the complete state (populations in their storage
 * precision, potential, macroscopic fields, step and time).  The reference's
 * restart file data_end.dat (LBM.cu:1505-1566, read back by read_data,
 * LBM.cu:1596-1643) keeps 6 decimals of the macroscopic fields and
 * re-equilibrates the populations; this one restarts bit for bit.  Loading
 * needs a box created with the same ABI version, grid, precision and physics
 * (T, C, M, Fe, both magic parameters, CFL, voltage, eps, rho0, Lz, uw, exf:
 * FP32S populations are stored against hydrostatic tables built from them);
 * any difference is EHD_EINVAL and the box is left unchanged. */
int ehd_save_state(ehd_box* b, const char* path);
int ehd_load_state(ehd_box* b, const char* path);

/* FP64 reductions on the device (current, total charge, umax, ...). */
int ehd_diagnostics(ehd_box* b, ehd_diag* out);

/* Reference current of the Nusselt number Nu = I / I_hs.  ehd_init sets
 * I_hs to ehd_hydrostatic_current of the box's parameters, so a box started
 * from the hydrostatic state has Nu = 1 at step 0; this call overrides it
 * (e.g. with the current of a converged 1-D run). */
int ehd_set_nusselt_reference(ehd_box* b, double current_hs);

/* Collector current of the analytic hydrostatic state sampled on the grid,
 * evaluated as current() (LBM.cu:1646-1671) evaluates it.  Host only.  The
 * lattice's own steady no-flow state differs from it at first order in dz
 * (13% at nz=17, 3.6% at nz=65 for C=10); ehd_hydrostatic_reference gives
 * that discrete value. */
int ehd_hydrostatic_current(const ehd_params* p, double* current);

/* Discrete hydrostatic current of the box: runs a 2x2 column of the same nz
 * and parameters (FP64, no perturbation) until the current changes by less
 * than tol (relative) over check_every steps, scales it to the box's nx*ny
 * columns, stores it as the box's Nusselt reference and returns it.  A box
 * below the convection threshold then settles at Nu = 1. */
int ehd_hydrostatic_reference(ehd_box* b, int check_every, double tol, long long max_steps, double* current);

/* Steady-state detector: steps the box in chunks of check_every steps and
 * stops once |Nu(after chunk) - Nu(before chunk)| < tol (e.g. 1e-6 over 1000
 * steps), or after max_steps.  steps counts the steps taken by this call.
 * A box the watchdog finds non-finite stops with failed = 1 and the call
 * returns EHD_ENONFINITE (out filled). */
typedef struct ehd_steady {
    long long steps;
    int converged;
    int failed;           /* the watchdog found non-finite fields (steps: the step of that check) */
    double nusselt, current, delta;
} ehd_steady;
int ehd_run_to_steady(ehd_box* b, int check_every, double tol, long long max_steps, ehd_steady* out);

/* Per-kernel timing with HIP events on the box's stream.  ehd_profile_enable
 * (b, n): n = 0 disables; n >= 1 brackets the fused LBM launch and the
 * Poisson chain of every step whose index is a multiple of n (n = 1: every
 * step; every launch of either does the same work, and each event record
 * idles the stream for a few microseconds, so a sample costs less), and
 * every ehd_step call; ehd_profile_read returns accumulated milliseconds and
 * counts. */
typedef struct ehd_profile {
    double lbm_ms;      long long lbm_launches;
    double poisson_ms;  long long poisson_calls;
    double step_ms;     long long steps;   /* ehd_step calls: their event-timed wall time and their number */
} ehd_profile;
int ehd_profile_enable(ehd_box* b, int on);
int ehd_profile_read(ehd_box* b, ehd_profile* out);
int ehd_profile_reset(ehd_box* b);

/* Debug mode (on != 0): every kernel launch of the box's steps and Poisson
 * solves is followed by a synchronisation of the box's stream and an error
 * check, so that an asynchronous fault is reported by the launch that caused
 * it, named in ehd_last_error() (phase and step).  The reference checks
 * launches only with cudaGetLastError and no synchronisation (LBM.cu:36-57),
 * so its faults surface later.  Off by default; the results are the same
 * either way, only slower. */
int ehd_set_debug_sync(ehd_box* b, int on);

int ehd_sync(ehd_box* b);
const char* ehd_last_error(void);

/* Practical HBM ceiling of the current GPU (no reference counterpart; the
 * reference computes a bandwidth figure at main.cu:251-260 but has no
 * ceiling to hold it against): a float4 device-to-device copy of `bytes`
 * bytes (bytes read and bytes written per launch), plain and with
 * nontemporal loads and stores, timed with HIP events over `reps` launches
 * after three warm-up launches.  out2[0] / out2[1]: GB/s (2 * bytes per
 * launch) of the plain / nontemporal copy.  Allocates two buffers of
 * `bytes` for the duration of the call. */
int ehd_copy_ceiling(size_t bytes, int reps, double* out2);

/* ---- the Poisson solve without allocation ----------------------------------
 * No counterpart in the reference: a variant of ehd_poisson for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It solves the FP64 Poisson
 * problem of ehd_poisson -- the same kernels, twiddles and job -- for the
 * grid and physics of *p (precision must be EHD_FP64) on caller-owned device
 * buffers: q (S doubles, [z][y][x], the charge) in, phi (S doubles, walls
 * included) out, on `stream` (a hipStream_t; NULL for the null stream).  Its
 * workspace (job table, twiddles, spectrum) is one file-scope device array of
 * EHD_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a grid
 * whose workspace (16 (nx/2+1) ny (nz-2) bytes of spectrum plus the twiddle
 * tables) does not fit is EHD_EINVAL (256x256x128 fits).  It calls no
 * allocating routine.  Every call on one GPU shares that array: calls that
 * overlap on one GPU race.  Returns after the stream has finished the solve. */
#define EHD_NOALLOC_SCRATCH_BYTES ((size_t)(64u << 20) + (size_t)(64u << 10))
int ehd_poisson_noalloc(const ehd_params* p, const double* q, double* phi, void* stream);

/* ---- batches: many independent boxes stepped together ------------------
 * Stability maps and hysteresis sweeps run many small boxes of uneven size
 * and parameters (T, C, M, lateral wavelength via nx, ny).  A batch steps
 * them together: one persistent launch of the fused LBM kernel over every
 * box (FP32S; FP64 boxes get one launch each), the Poisson solve batched
 * per (nx, ny) group (one launch per FFT / z-solve pass for all boxes of the
 * group), and the steady-state part of the step sequence captured once in a
 * HIP graph and replayed.  Each box gives the same numbers, bit for bit, as
 * when stepped alone.  No reference counterpart: the reference runs one box
 * per process (main.cu). */
typedef struct ehd_batch ehd_batch;

#define EHD_BATCH_GRAPH 1   /* replay the steady steps from a captured HIP graph */
#define EHD_BATCH_TUNE  2   /* time the launch shapes at creation and keep the fastest */
#define EHD_BATCH_CHAINS 4  /* one LBM launch and Poisson chain per (nx, ny) group, the groups on
                               concurrent streams (with EHD_BATCH_TUNE and without this flag the
                               tuner times the steady step both ways, replayed from a graph when
                               EHD_BATCH_GRAPH is set, and keeps the faster) */

/* Launch shapes of a batch (the autotuner's choice or the defaults). */
typedef struct ehd_tuning {
    int lbm_grid;             /* workgroups of the persistent LBM launch (= tiles when not persistent) */
    int lbm_tiles;            /* 256-node tiles over all boxes */
    double lbm_ms;            /* timed LBM launch (tuning run), 0 if not tuned */
    int ngroups;              /* Poisson groups (distinct nx, ny) */
    struct {
        int nx, ny, nboxes, x_lines, y_lines;   /* lines per workgroup of the x / y FFT passes */
        double ms;                               /* timed Poisson chain of the group, 0 if not tuned */
    } group[16];
    int lbm_chunk;            /* XCD tile order of the LBM launch: tiles per XCD run (0: linear order) */
    int chains;               /* 1: each group runs its own LBM launch and Poisson chain on its stream */
    double step_ms[2];        /* timed step with one LBM launch / with chains (tuning), 0 if not timed */
} ehd_tuning;

/* nboxes parameter sets (all of one precision; at most 16 distinct (nx, ny)). */
int ehd_batch_create(const ehd_params* params, int nboxes, int flags, ehd_batch** out);
void ehd_batch_destroy(ehd_batch* bt);
/* ehd_init of every box; fields9[i] as ehd_init's argument (fields9 itself may be NULL). */
int ehd_batch_init(ehd_batch* bt, const double* const* fields9);
int ehd_batch_step(ehd_batch* bt, int nsteps);
/* ehd_init_consistent for the boxes with fields9[i] != NULL (the others start
 * as ehd_init does and step normally meanwhile); every box advances
 * `iterations` steps.  FP32S batches. */
int ehd_batch_init_consistent(ehd_batch* bt, const double* const* fields9, int iterations);
/* ehd_diagnostics of every box (out: nboxes entries, all filled); returns
 * EHD_ENONFINITE if the watchdog flagged any box (out[i].nonfinite > 0). */
int ehd_batch_diagnostics(ehd_batch* bt, ehd_diag* out);
/* steps the batch in chunks of check_every until every box has |dNu| < tol
 * over a chunk (or max_steps); out[i].steps is the step count at which box i
 * first met the criterion (max_steps if never).  A box the watchdog flags
 * gets failed = 1 (steps: the check that found it) and is no longer tested;
 * the batch keeps stepping the others (a box's non-finite values stay in its
 * own buffers), and the call returns EHD_OK. */
int ehd_batch_run_to_steady(ehd_batch* bt, int check_every, double tol, long long max_steps, ehd_steady* out);
/* member box i (owned by the batch; valid until ehd_batch_destroy) for the
 * per-box calls: fields, populations, checkpoints, Nusselt reference. */
ehd_box* ehd_batch_box(ehd_batch* bt, int i);
int ehd_batch_tuning(ehd_batch* bt, ehd_tuning* out);
int ehd_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
