/*
 * ehd_oracle.c -- FP64 CPU oracle for the coupled EHD time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function restates the reference kernel named in its comment, in the
 * reference's arithmetic order (the association of every sum and product is
 * kept, so that an implementation following the same order can be compared
 * bit for bit).  Runtime grid sizes replace LBM.h's compile-time constants.
 * Build with -ffp-contract=off: no fused multiply-adds anywhere.
 *
 * Known reference race, resolved here: the z=0 velocity override of
 * gpu_collide_save (LBM.cu:416-493) reads h0 at z=1 (LBM.cu:419) while the
 * z=1 thread of the same launch updates h0 in place (LBM.cu:1031).  The z=0
 * blocks are dispatched first and read it before the z=1 blocks write it, so
 * the oracle takes the start-of-step (pre-collision) value.
 */
#include "ehd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------- */
/* Lattice tables.                                                            */
/* ------------------------------------------------------------------------- */

/* Direction vectors, read off the pull offsets of gpu_stream (LBM.cu:1225-1250):
 * f1[d](x) = f2[d](x - c_d). */
static const int CV[27][3] = {
    { 0, 0, 0},
    { 1, 0, 0}, {-1, 0, 0}, { 0, 1, 0}, { 0,-1, 0}, { 0, 0, 1}, { 0, 0,-1},
    { 1, 1, 0}, {-1,-1, 0}, { 1, 0, 1}, {-1, 0,-1}, { 0, 1, 1}, { 0,-1,-1},
    { 1,-1, 0}, {-1, 1, 0}, { 1, 0,-1}, {-1, 0, 1}, { 0, 1,-1}, { 0,-1, 1},
    { 1, 1, 1}, {-1,-1,-1}, { 1, 1,-1}, {-1,-1, 1}, { 1,-1, 1}, {-1, 1,-1},
    {-1, 1, 1}, { 1,-1,-1}};

/* Signed axis terms (1=x, 2=y, 3=z) of c.u, in the order the reference
 * evaluates them: flow equilibria LBM.cu:545-660 (and 169-223), charge
 * equilibria LBM.cu:560-660 (and 178-272).  They differ only for d=11 and 20,
 * whose values coincide exactly. */
static const int EQF[27][3] = {
    {0,0,0},
    { 1,0,0}, {-1,0,0}, { 2,0,0}, {-2,0,0}, { 3,0,0}, {-3,0,0},
    { 1, 2,0}, {-2,-1,0}, { 1, 3,0}, {-1,-3,0}, { 3, 2,0}, {-2,-3,0},
    { 1,-2,0}, { 2,-1,0}, { 1,-3,0}, { 3,-1,0}, { 2,-3,0}, { 3,-2,0},
    { 1, 2, 3}, {-2,-1,-3}, { 1, 2,-3}, { 3,-1,-2}, { 1, 3,-2}, { 2,-1,-3},
    { 2, 3,-1}, { 1,-2,-3}};
static const int EQH[27][3] = {
    {0,0,0},
    { 1,0,0}, {-1,0,0}, { 2,0,0}, {-2,0,0}, { 3,0,0}, {-3,0,0},
    { 1, 2,0}, {-2,-1,0}, { 1, 3,0}, {-1,-3,0}, { 2, 3,0}, {-2,-3,0},
    { 1,-2,0}, { 2,-1,0}, { 1,-3,0}, { 3,-1,0}, { 2,-3,0}, { 3,-2,0},
    { 1, 2, 3}, {-1,-2,-3}, { 1, 2,-3}, { 3,-1,-2}, { 1, 3,-2}, { 2,-1,-3},
    { 2, 3,-1}, { 1,-2,-3}};

static int opp_of(int d) { return d == 0 ? 0 : ((d & 1) ? d + 1 : d - 1); }

static double weight_of(const oref_params* p, int d)
{
    if (d == 0) return p->w0;
    if (d <= 6) return p->ws;
    if (d <= 18) return p->wa;
    return p->wd;
}

/* Momentum sums of LBM.cu:382-387: directions with c_a=+1 in ascending order,
 * minus those with c_a=-1 in ascending order. */
static int MOM[3][18];
static void build_mom(void)
{
    for (int a = 0; a < 3; ++a) {
        int np = 0, nn = 9;
        for (int d = 1; d < 27; ++d) {
            if (CV[d][a] == 1) MOM[a][np++] = d;
            else if (CV[d][a] == -1) MOM[a][nn++] = d;
        }
    }
}

void oref_tables(int* c, int* opp, double* w, int* eqf, int* eqh, int* mom)
{
    oref_params p;
    p.w0 = 8.0 / 27.0; p.ws = 2.0 / 27.0; p.wa = 1.0 / 54.0; p.wd = 1.0 / 216.0;
    build_mom();
    for (int d = 0; d < 27; ++d) {
        for (int a = 0; a < 3; ++a) c[3 * d + a] = CV[d][a];
        opp[d] = opp_of(d);
        w[d] = weight_of(&p, d);
        for (int k = 0; k < 3; ++k) { eqf[4 * d + k] = EQF[d][k]; eqh[4 * d + k] = EQH[d][k]; }
        eqf[4 * d + 3] = 0; eqh[4 * d + 3] = 0;
    }
    for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 18; ++k) mom[18 * a + k] = MOM[a][k];
}

void oref_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------- */
/* Parameters.                                                                */
/* ------------------------------------------------------------------------- */

int oref_params_init(oref_params* p, int nx, int ny, int nz,
                     double T, double M, double C, double Fe,
                     double lambda_f, double lambda_c,
                     double cfl, double voltage, double eps, double rho0, double lz,
                     double uw, double exf)
{
    if (!p || nx < 2 || ny < 2 || nz < 4) return -1;
    memset(p, 0, sizeof(*p));
    p->nx = nx; p->ny = ny; p->nz = nz; p->ne = 2 * (nz - 1);
    p->lz = lz;
    p->dz = lz / (double)(nz - 1);
    p->dx = p->dz; p->dy = p->dz;
    p->lx = nx * p->dx; p->ly = ny * p->dy;
    p->cfl = cfl;
    p->dt = cfl * p->dx;                              /* LBM.h:52 dt = CFL*dx */
    p->cs_square = 1.0 / 3.0 / (cfl * cfl);           /* LBM.h:53 */
    p->rho0 = rho0; p->voltage = voltage; p->eps = eps;
    p->K = sqrt(eps / rho0) / M;
    p->nu = eps * voltage / (p->K * rho0 * T);
    p->charge0 = C * voltage * eps / (lz * lz);
    p->diffu = p->K * voltage / Fe;
    p->uw = uw; p->exf = exf;
    p->V = lambda_f; p->VC = lambda_c;
    p->w0 = 8.0 / 27.0; p->ws = 2.0 / 27.0; p->wa = 1.0 / 54.0; p->wd = 1.0 / 216.0;
    return 0;
}

void oref_compute_parameters(const oref_params* p, double* T, double* M, double* C, double* Fe)
{
    /* LBM.cu:1375-1378 */
    *M = sqrt(p->eps / p->rho0) / p->K;
    *T = p->eps * p->voltage / p->K / p->nu / p->rho0;
    *C = p->charge0 * p->lz * p->lz / (p->voltage * p->eps);
    *Fe = p->K * p->voltage / p->diffu;
}

/* ------------------------------------------------------------------------- */
/* FFT: unnormalised DFT, sign -1 forward (cuFFT CUFFT_FORWARD), +1 inverse.  */
/* Mixed radix (2,3,4,5,7 and small odd factors) with Bluestein for lengths  */
/* holding a prime factor above 13.                                           */
/* ------------------------------------------------------------------------- */

typedef struct fftplan {
    int n;
    int nf;
    int fac[40];
    double* tw;          /* 2n: cos, sin of 2*pi*j/n */
    int blue;            /* Bluestein */
    int m;               /* Bluestein convolution length (power of 2) */
    double* chirp;       /* 2n: exp(i*pi*(j^2 mod 2n)/n) */
    double* bfft[2];     /* 2m: DFT of conj(chirp) (sign-1) / chirp (sign+1), index by (sign>0) */
    struct fftplan* sub; /* plan of length m */
} fftplan;

static void plan_free(fftplan* P)
{
    if (!P) return;
    free(P->tw); free(P->chirp); free(P->bfft[0]); free(P->bfft[1]);
    if (P->sub) plan_free(P->sub);
    free(P);
}

static void fft_exec(const fftplan* P, double* data, int sign, double* work);

static fftplan* plan_make(int n)
{
    fftplan* P = (fftplan*)calloc(1, sizeof(fftplan));
    P->n = n;
    P->tw = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (int j = 0; j < n; ++j) {
        double a = 2.0 * M_PI * (double)j / (double)n;
        P->tw[2 * j] = cos(a);
        P->tw[2 * j + 1] = sin(a);
    }
    int r = n, maxp = 1;
    static const int pref[] = {4, 2, 3, 5, 7};
    for (int i = 0; i < 5; ++i)
        while (r % pref[i] == 0) { P->fac[P->nf++] = pref[i]; r /= pref[i]; if (pref[i] > maxp) maxp = pref[i]; }
    for (int q = 11; r > 1; q += 2)
        while (r % q == 0) { P->fac[P->nf++] = q; r /= q; if (q > maxp) maxp = q; }
    if (maxp > 13) {
        P->blue = 1;
        int m = 1;
        while (m < 2 * n - 1) m <<= 1;
        P->m = m;
        P->sub = plan_make(m);
        P->chirp = (double*)malloc(sizeof(double) * 2 * (size_t)n);
        for (int j = 0; j < n; ++j) {
            long long jj = ((long long)j * j) % (2LL * n);
            double a = M_PI * (double)jj / (double)n;
            P->chirp[2 * j] = cos(a);
            P->chirp[2 * j + 1] = sin(a);
        }
        double* work = (double*)malloc(sizeof(double) * 2 * (size_t)m);
        for (int s = 0; s < 2; ++s) {
            /* sign -1 (s=0): X_k = w_k * sum_j (x_j w_j) conj(w_{k-j}) with w = conj(chirp);
             * b_j = conj(w_j) = chirp_j.  sign +1 (s=1): w = chirp, b_j = conj(chirp_j). */
            double* b = (double*)calloc(2 * (size_t)m, sizeof(double));
            double sg = (s == 0) ? 1.0 : -1.0;
            b[0] = P->chirp[0]; b[1] = sg * P->chirp[1];
            for (int j = 1; j < n; ++j) {
                b[2 * j] = P->chirp[2 * j]; b[2 * j + 1] = sg * P->chirp[2 * j + 1];
                b[2 * (m - j)] = P->chirp[2 * j]; b[2 * (m - j) + 1] = sg * P->chirp[2 * j + 1];
            }
            fft_exec(P->sub, b, -1, work);
            P->bfft[s] = b;
        }
        free(work);
    }
    return P;
}

/* Recursive decimation in time: out[k] = sum_j in[j*istride] W_n^{s jk}. */
static void fft_rec(const fftplan* P, const double* in, double* out, int n, int istride, int fi, int sign)
{
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int p = P->fac[fi];
    int m = n / p;
    for (int r = 0; r < p; ++r)
        fft_rec(P, in + 2 * (size_t)r * istride, out + 2 * (size_t)r * m, m, istride * p, fi + 1, sign);
    const int tstep = P->n / n;          /* W_n^j = tw[j*tstep]; r*k*tstep < P->n */
    const double* tw = P->tw;
    const double sg = (double)sign;
#define TWR(j) (tw[2 * (size_t)(j)])
#define TWI(j) (sg * tw[2 * (size_t)(j) + 1])
    if (p == 2) {
        for (int k = 0; k < m; ++k) {
            double* A = out + 2 * (size_t)k;
            double* B = out + 2 * ((size_t)m + k);
            int j = k * tstep;
            double br = B[0] * TWR(j) - B[1] * TWI(j), bi = B[0] * TWI(j) + B[1] * TWR(j);
            double ar = A[0], ai = A[1];
            A[0] = ar + br; A[1] = ai + bi;
            B[0] = ar - br; B[1] = ai - bi;
        }
        return;
    }
    if (p == 4) {
        for (int k = 0; k < m; ++k) {
            double* X0 = out + 2 * (size_t)k;
            double* X1 = out + 2 * ((size_t)m + k);
            double* X2 = out + 2 * (2 * (size_t)m + k);
            double* X3 = out + 2 * (3 * (size_t)m + k);
            int j1 = k * tstep, j2 = 2 * k * tstep, j3 = 3 * k * tstep;
            double ar = X0[0], ai = X0[1];
            double br = X1[0] * TWR(j1) - X1[1] * TWI(j1), bi = X1[0] * TWI(j1) + X1[1] * TWR(j1);
            double cr = X2[0] * TWR(j2) - X2[1] * TWI(j2), ci = X2[0] * TWI(j2) + X2[1] * TWR(j2);
            double dr = X3[0] * TWR(j3) - X3[1] * TWI(j3), di = X3[0] * TWI(j3) + X3[1] * TWR(j3);
            double t0r = ar + cr, t0i = ai + ci, t1r = ar - cr, t1i = ai - ci;
            double t2r = br + dr, t2i = bi + di, t3r = br - dr, t3i = bi - di;
            double ir = -sg * t3i, ii = sg * t3r;     /* W4 = i*sign */
            X0[0] = t0r + t2r; X0[1] = t0i + t2i;
            X1[0] = t1r + ir;  X1[1] = t1i + ii;
            X2[0] = t0r - t2r; X2[1] = t0i - t2i;
            X3[0] = t1r - ir;  X3[1] = t1i - ii;
        }
        return;
    }
    double t[2 * 64], wp[2 * 64];
    for (int j = 0; j < p; ++j) { wp[2 * j] = TWR((size_t)j * (P->n / p)); wp[2 * j + 1] = TWI((size_t)j * (P->n / p)); }
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) {
            double xr = out[2 * ((size_t)r * m + k)], xi = out[2 * ((size_t)r * m + k) + 1];
            int j = r * k * tstep;
            t[2 * r] = xr * TWR(j) - xi * TWI(j);
            t[2 * r + 1] = xr * TWI(j) + xi * TWR(j);
        }
        for (int q = 0; q < p; ++q) {
            double ar = 0.0, ai = 0.0;
            int idx = 0;
            for (int r = 0; r < p; ++r) {
                ar += t[2 * r] * wp[2 * idx] - t[2 * r + 1] * wp[2 * idx + 1];
                ai += t[2 * r] * wp[2 * idx + 1] + t[2 * r + 1] * wp[2 * idx];
                idx += q;
                if (idx >= p) idx -= p;
            }
            out[2 * ((size_t)q * m + k)] = ar;
            out[2 * ((size_t)q * m + k) + 1] = ai;
        }
    }
#undef TWR
#undef TWI
}

/* data: 2n doubles, transformed in place.  work: at least 2*max(n, 2m) doubles. */
static void fft_exec(const fftplan* P, double* data, int sign, double* work)
{
    int n = P->n;
    if (!P->blue) {
        fft_rec(P, data, work, n, 1, 0, sign);
        memcpy(data, work, sizeof(double) * 2 * (size_t)n);
        return;
    }
    int m = P->m;
    double* a = work;              /* 2m */
    double* w2 = work + 2 * (size_t)m;   /* 2m scratch for the sub FFT */
    memset(a, 0, sizeof(double) * 2 * (size_t)m);
    double sg = (sign < 0) ? -1.0 : 1.0;   /* w_j = chirp^sg */
    for (int j = 0; j < n; ++j) {
        double wr = P->chirp[2 * j], wi = sg * P->chirp[2 * j + 1];
        double xr = data[2 * j], xi = data[2 * j + 1];
        a[2 * j] = xr * wr - xi * wi;
        a[2 * j + 1] = xr * wi + xi * wr;
    }
    fft_exec(P->sub, a, -1, w2);
    const double* b = P->bfft[sign < 0 ? 0 : 1];
    for (int k = 0; k < m; ++k) {
        double ar = a[2 * k], ai = a[2 * k + 1];
        a[2 * k] = ar * b[2 * k] - ai * b[2 * k + 1];
        a[2 * k + 1] = ar * b[2 * k + 1] + ai * b[2 * k];
    }
    fft_exec(P->sub, a, +1, w2);
    double inv = 1.0 / (double)m;
    for (int k = 0; k < n; ++k) {
        double wr = P->chirp[2 * k], wi = sg * P->chirp[2 * k + 1];
        double cr = a[2 * k] * inv, ci = a[2 * k + 1] * inv;
        data[2 * k] = cr * wr - ci * wi;
        data[2 * k + 1] = cr * wi + ci * wr;
    }
}

static size_t plan_work(const fftplan* P)
{
    if (!P->blue) return 2 * (size_t)P->n;
    return 4 * (size_t)P->m + plan_work(P->sub);
}

void oref_fft1d(double* data, int n, int sign)
{
    fftplan* P = plan_make(n);
    double* work = (double*)malloc(sizeof(double) * (plan_work(P) + 2 * (size_t)n));
    fft_exec(P, data, sign, work);
    free(work);
    plan_free(P);
}

/* Transform along one axis of a [n2][n1][n0] complex array (axis 0 fastest). */
static void fft_axis(const fftplan* P, double* a, int n0, int n1, int n2, int axis, int sign)
{
    size_t nlines, stride;
    int n;
    if (axis == 0) { n = n0; stride = 1; nlines = (size_t)n1 * n2; }
    else if (axis == 1) { n = n1; stride = (size_t)n0; nlines = (size_t)n0 * n2; }
    else { n = n2; stride = (size_t)n0 * n1; nlines = (size_t)n0 * n1; }
    size_t wsz = plan_work(P) + 2 * (size_t)n;
#pragma omp parallel
    {
        double* line = (double*)malloc(sizeof(double) * (2 * (size_t)n + wsz));
        double* work = line + 2 * (size_t)n;
#pragma omp for schedule(static)
        for (long long L = 0; L < (long long)nlines; ++L) {
            size_t base;
            if (axis == 0) base = (size_t)L * n0;
            else if (axis == 1) base = ((size_t)L / n0) * (size_t)n0 * n1 + ((size_t)L % n0);
            else base = (size_t)L;
            for (int j = 0; j < n; ++j) {
                line[2 * j] = a[2 * (base + j * stride)];
                line[2 * j + 1] = a[2 * (base + j * stride) + 1];
            }
            fft_exec(P, line, sign, work);
            for (int j = 0; j < n; ++j) {
                a[2 * (base + j * stride)] = line[2 * j];
                a[2 * (base + j * stride) + 1] = line[2 * j + 1];
            }
        }
        free(line);
    }
}

/* ------------------------------------------------------------------------- */
/* Box.                                                                       */
/* ------------------------------------------------------------------------- */

struct oref_box {
    oref_params p;
    size_t S, NXY, SE;
    double* fld[OREF_F_COUNT];
    fftplan *px, *py, *pz;
    /* constants of gpu_collide_save (LBM.cu:298-302) */
    double omega_plus, omega_minus, omega_c_minus, omega_c_plus;
};

oref_box* oref_create(const oref_params* p)
{
    oref_box* b = (oref_box*)calloc(1, sizeof(oref_box));
    b->p = *p;
    build_mom();
    size_t S = (size_t)p->nx * p->ny * p->nz;
    b->S = S; b->NXY = (size_t)p->nx * p->ny; b->SE = (size_t)p->nx * p->ny * p->ne;
    size_t sz[OREF_F_COUNT] = {
        S, 26 * S, 26 * S, S, 26 * S, 26 * S,
        S, S, S, S, S, S, S, S, S, 2 * b->NXY,
        2 * b->SE, 2 * b->SE, 2 * b->SE, (size_t)p->nx, (size_t)p->ny, (size_t)p->ne};
    for (int i = 0; i < OREF_F_COUNT; ++i) b->fld[i] = (double*)calloc(sz[i], sizeof(double));
    b->px = plan_make(p->nx); b->py = plan_make(p->ny); b->pz = plan_make(p->ne);

    /* k-vectors, main.cu:109-131 */
    double* kx = b->fld[OREF_F_KX]; double* ky = b->fld[OREF_F_KY]; double* kz = b->fld[OREF_F_KZ];
    unsigned NX = p->nx, NY = p->ny, NE = p->ne;
    for (unsigned i = 0; i <= NX / 2; i++) kx[i] = (double)i * 2.0 * M_PI / p->lx;
    for (unsigned i = NX / 2 + 1; i < NX; i++) kx[i] = ((double)i - NX) * 2.0 * M_PI / p->lx;
    for (unsigned i = 0; i <= NY / 2; i++) ky[i] = (double)i * 2.0 * M_PI / p->ly;
    for (unsigned i = NY / 2 + 1; i < NY; i++) ky[i] = ((double)i - NY) * 2.0 * M_PI / p->ly;
    for (unsigned i = 0; i <= NE / 2; i++) kz[i] = (double)i * 2.0 * M_PI / (NE * p->dz);
    for (unsigned i = NE / 2 + 1; i < NE; i++) kz[i] = ((double)i - NE) * 2.0 * M_PI / (NE * p->dz);

    /* LBM.cu:298-302 */
    double nu = p->nu, cs = p->cs_square, dt = p->dt, diffu = p->diffu;
    b->omega_plus = 1.0 / (nu / cs / dt + 1.0 / 2.0) / dt;
    b->omega_minus = 1.0 / (p->V / (nu / cs / dt) + 1.0 / 2.0) / dt;
    b->omega_c_minus = 1.0 / (diffu / cs / dt + 1.0 / 2.0) / dt;
    b->omega_c_plus = 1.0 / (p->VC / (diffu / cs / dt) + 1.0 / 2.0) / dt;
    return b;
}

void oref_destroy(oref_box* b)
{
    if (!b) return;
    for (int i = 0; i < OREF_F_COUNT; ++i) free(b->fld[i]);
    plan_free(b->px); plan_free(b->py); plan_free(b->pz);
    free(b);
}

const oref_params* oref_get_params(const oref_box* b) { return &b->p; }
double* oref_field(oref_box* b, int which)
{
    return (which >= 0 && which < OREF_F_COUNT) ? b->fld[which] : NULL;
}

#define SIDX(x, y, z) ((size_t)NX * ((size_t)NY * (z) + (y)) + (x))
#define NIDX(x, y, z, d) ((size_t)NX * ((size_t)NY * ((size_t)NZ * ((d) - 1) + (z)) + (y)) + (x))

/* ------------------------------------------------------------------------- */
/* Initial states.                                                            */
/* ------------------------------------------------------------------------- */

void oref_initialization(oref_box* b)
{
    /* gpu_initialization, LBM.cu:80-96 */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz;
    for (unsigned z = 0; z < NZ; ++z)
        for (unsigned y = 0; y < NY; ++y)
            for (unsigned x = 0; x < NX; ++x) {
                size_t s = SIDX(x, y, z);
                b->fld[OREF_F_RHO][s] = p->rho0;
                b->fld[OREF_F_Q][s] = 0.0;
                b->fld[OREF_F_PHI][s] = p->voltage * (p->lz - p->dz * z) / p->lz;
                b->fld[OREF_F_UX][s] = 0.0; b->fld[OREF_F_UY][s] = 0.0; b->fld[OREF_F_UZ][s] = 0.0;
                b->fld[OREF_F_EX][s] = 0.0; b->fld[OREF_F_EY][s] = 0.0; b->fld[OREF_F_EZ][s] = 0.0;
            }
}

void oref_set_fields(oref_box* b, const double* f9)
{
    static const int order[9] = {OREF_F_RHO, OREF_F_Q, OREF_F_PHI, OREF_F_UX, OREF_F_UY, OREF_F_UZ,
                                 OREF_F_EX, OREF_F_EY, OREF_F_EZ};
    for (int i = 0; i < 9; ++i) memcpy(b->fld[order[i]], f9 + (size_t)i * b->S, sizeof(double) * b->S);
}

static uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double oref_hash_uniform(uint64_t seed, uint32_t x, uint32_t y, uint32_t z, uint32_t comp)
{
    uint64_t key = ((uint64_t)comp << 60) ^ ((uint64_t)z << 40) ^ ((uint64_t)y << 20) ^ (uint64_t)x;
    uint64_t h = mix64(seed ^ mix64(key));
    double u = (double)(h >> 11) * 0x1.0p-53;
    return 2.0 * u - 1.0;
}

/* Nondimensional hydrostatic state with D=0, u=0: E^2 = E0^2 + 2 C E0 z,
 * q = C E0 / E, phi = 1 - (E^3 - E0^3)/(3 C E0), with phi(1) = 0. */
double oref_hydrostatic_e0(double C)
{
    /* g(e0) = (E(1)^3 - e0^3)/(3 C e0) - 1 rises from -1 (e0 -> 0) through 0;
     * bisection on [0, 1] to the last bit. */
    double lo = 0.0, hi = 1.0;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        double e1 = sqrt(mid * mid + 2.0 * C * mid);
        double g = (e1 * e1 * e1 - mid * mid * mid) / (3.0 * C * mid) - 1.0;
        if (g > 0.0) hi = mid; else lo = mid;
    }
    return 0.5 * (lo + hi);
}

/* Velocity perturbation patterns of gpu_collide_save's perturb branch
 * (LBM.cu:389-411): pattern 1 rolls (LBM.cu:394-395), 2 squares
 * (LBM.cu:397-399), 3 crossed rolls (LBM.cu:400-402, the active form), 4
 * hexagons (LBM.cu:404-411).  The reference's wavelength LL (LBM.h:39, 1.22)
 * is the box width Lx of its 122-node grid; here the wavelength is the box
 * width Lx (Ly for y), and the hexagon scale L = Lx/sqrt(3). */
void oref_pattern_velocity(int pattern, double amp, double xx, double yy, double zz, double lx, double ly,
                           double lz, double* u)
{
    const double pi = 3.14159265358979323846;
    double zeta = zz / lz;
    double cz = cos(2.0 * pi * zeta) - 1.0, sz = sin(2.0 * pi * zeta);
    u[0] = u[1] = u[2] = 0.0;
    switch (pattern) {
    case 1:
        u[2] = amp * cz * cos(2.0 * pi / lx * xx);
        u[0] = amp * (lx / lz) * sz * sin(2.0 * pi / lx * xx);
        break;
    case 2:
        u[2] = amp * cz * cos(2.0 * pi / lx * xx) * cos(2.0 * pi / ly * yy);
        u[0] = 0.5 * amp * (lx / lz) * sz * sin(2.0 * pi / lx * xx) * cos(2.0 * pi / ly * yy);
        u[1] = 0.5 * amp * (ly / lz) * sz * sin(2.0 * pi / ly * yy) * cos(2.0 * pi / lx * xx);
        break;
    case 3:
        u[2] = amp * cz * (cos(2.0 * pi / lx * xx) + cos(2.0 * pi / ly * yy));
        u[0] = amp * (lx / lz) * sz * sin(2.0 * pi / lx * xx);
        u[1] = amp * (ly / lz) * sz * sin(2.0 * pi / ly * yy);
        break;
    case 4: {
        double L = lx / sqrt(3.0), a = 4.0 * pi / 3.0 / L;
        double kx = 2.0 * pi / (sqrt(3.0) * L), ky = 2.0 * pi / (3.0 * L);
        u[2] = -amp * cz / 3.0 * (2.0 * cos(kx * xx) * cos(ky * yy) + cos(2.0 * ky * yy));
        u[0] = -amp * 2.0 * pi * sz * 4.0 * pi / (3.0 * sqrt(3.0) * L * a * a) * sin(kx * xx) * cos(ky * yy);
        u[1] = -amp * 2.0 * pi * sz * 4.0 * pi / (9.0 * L * a * a) * (cos(kx * xx) + 2.0 * cos(ky * yy))
               * sin(ky * yy);
        break;
    }
    default:
        break;
    }
}

void oref_init_hydrostatic(oref_box* b, uint64_t seed, double amp, int hydro_density)
{
    oref_init_hydrostatic_pattern(b, seed, amp, hydro_density, 0);
}

void oref_init_hydrostatic_pattern(oref_box* b, uint64_t seed, double amp, int hydro_density, int pattern)
{
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz;
    double C = p->charge0 * p->lz * p->lz / (p->voltage * p->eps);
    double e0 = oref_hydrostatic_e0(C);
    double qs = p->eps * p->voltage / (p->lz * p->lz);
    /* uniform current density J = q Ez balances the pressure: dp/dz = J,
     * rho = rho0 + J (z - Lz/2)/cs^2 (mean rho0) */
    double J = C * e0 * p->eps * p->voltage * p->voltage / (p->lz * p->lz * p->lz);
    for (unsigned z = 0; z < NZ; ++z) {
        double zeta = p->dz * z / p->lz;
        double e = sqrt(e0 * e0 + 2.0 * C * e0 * zeta);
        double qz = C * e0 / e * qs;
        double phz = p->voltage * (1.0 - (e * e * e - e0 * e0 * e0) / (3.0 * C * e0));
        double rhoz = hydro_density ? p->rho0 + J * (p->dz * z - 0.5 * p->lz) / p->cs_square : p->rho0;
        if (z == 0) phz = p->voltage;
        if (z == NZ - 1) phz = 0.0;
        for (unsigned y = 0; y < NY; ++y)
            for (unsigned x = 0; x < NX; ++x) {
                size_t s = SIDX(x, y, z);
                b->fld[OREF_F_RHO][s] = rhoz;
                b->fld[OREF_F_Q][s] = qz;
                b->fld[OREF_F_PHI][s] = phz;
                int wall = (z == 0 || z == NZ - 1);
                double u[3];
                if (pattern == 0) {
                    u[0] = amp * oref_hash_uniform(seed, x, y, z, 0);
                    u[1] = amp * oref_hash_uniform(seed, x, y, z, 1);
                    u[2] = amp * oref_hash_uniform(seed, x, y, z, 2);
                } else {
                    oref_pattern_velocity(pattern, amp, p->dx * x, p->dy * y, p->dz * z, p->lx, p->ly, p->lz, u);
                }
                b->fld[OREF_F_UX][s] = wall ? 0.0 : u[0];
                b->fld[OREF_F_UY][s] = wall ? 0.0 : u[1];
                b->fld[OREF_F_UZ][s] = wall ? 0.0 : u[2];
            }
    }
    oref_efield(b);
}

/* c.u of direction d from the signed axis terms, evaluated left to right. */
static inline double cdot_terms(const int* tm, const double* t /* t[1..3] */)
{
    int k = tm[0];
    double acc = (k > 0) ? t[k] : -t[-k];
    for (int i = 1; i < 3 && tm[i] != 0; ++i) {
        k = tm[i];
        acc = (k > 0) ? acc + t[k] : acc - t[-k];
    }
    return acc;
}

/* Equilibria of one node (LBM.cu:519-660; LBM.cu:140-272 for the init kernel).
 * fe[27]: flow with (rho,u); he[27]: charge with (q, u+K*E). */
static void node_equilibria(const oref_params* p, double rho, double ux, double uy, double uz,
                            double charge, double Ex, double Ey, double Ez,
                            double* fe, double* he)
{
    double cs_square = p->cs_square, CFL = p->cfl, K = p->K;
    double wr[4] = {p->w0 * rho, p->ws * rho, p->wa * rho, p->wd * rho};
    double wc[4] = {p->w0 * charge, p->ws * charge, p->wa * charge, p->wd * charge};
    double omusq = 1.0 - 0.5 * (ux * ux + uy * uy + uz * uz) / cs_square;
    double omusq_c = 1.0 - 0.5 * ((ux + K * Ex) * (ux + K * Ex) + (uy + K * Ey) * (uy + K * Ey)
                                  + (uz + K * Ez) * (uz + K * Ez)) / cs_square;
    double t[4], tc[4];
    t[1] = ux / cs_square / CFL; t[2] = uy / cs_square / CFL; t[3] = uz / cs_square / CFL;
    tc[1] = (ux + K * Ex) / cs_square / CFL;
    tc[2] = (uy + K * Ey) / cs_square / CFL;
    tc[3] = (uz + K * Ez) / cs_square / CFL;
    fe[0] = wr[0] * (omusq);
    he[0] = wc[0] * (omusq_c);
    for (int d = 1; d < 27; ++d) {
        int cls = (d <= 6) ? 1 : (d <= 18) ? 2 : 3;
        double cu = cdot_terms(EQF[d], t);
        fe[d] = wr[cls] * (omusq + cu * (1.0 + 0.5 * cu));
        double cc = cdot_terms(EQH[d], tc);
        he[d] = wc[cls] * (omusq_c + cc * (1.0 + 0.5 * cc));
    }
}

/* Guo force populations (LBM.cu:664-702).  Per direction with nonzero axes
 * a: A_a = (c_a*cflinv - u_a) + P_a*cflinv2, P_a = sum_b (c_a c_b) u_b over
 * the nonzero axes in x,y,z order; the directions with 1, 2 and 3 nonzero
 * axes combine the A_a F_a and u_b F_b terms in the reference's order. */
static void node_forcing(const oref_params* p, double ux, double uy, double uz,
                         double forcex, double forcey, double forcez, double* fpop)
{
    double cs_square = p->cs_square;
    double coe[4] = {p->w0 / cs_square, p->ws / cs_square, p->wa / cs_square, p->wd / cs_square};
    double cflinv = 1.0 / p->cfl;
    double cflinv2 = cflinv * cflinv / cs_square;
    double u[3] = {ux, uy, uz};
    double F[3] = {forcex, forcey, forcez};
    fpop[0] = -coe[0] * (ux * forcex + uy * forcey + uz * forcez);
    for (int d = 1; d < 27; ++d) {
        int ax[3], na = 0, zero[3], nz0 = 0;
        for (int a = 0; a < 3; ++a) { if (CV[d][a]) ax[na++] = a; else zero[nz0++] = a; }
        double A[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < na; ++i) {
            int a = ax[i];
            double P = 0.0;
            for (int j = 0; j < na; ++j) {
                int bb = ax[j];
                int s = CV[d][a] * CV[d][bb];
                if (j == 0) P = (s > 0) ? u[bb] : -u[bb];
                else P = (s > 0) ? P + u[bb] : P - u[bb];
            }
            double head = (CV[d][a] > 0) ? (cflinv - u[a]) : (-cflinv - u[a]);
            A[i] = head + P * cflinv2;
        }
        double sum;
        if (na == 1) {
            int b0 = zero[0], b1 = zero[1];
            sum = -u[b0] * F[b0] - u[b1] * F[b1] + A[0] * F[ax[0]];
        } else if (na == 2) {
            sum = A[0] * F[ax[0]] + A[1] * F[ax[1]] - u[zero[0]] * F[zero[0]];
        } else {
            sum = A[0] * F[0] + A[1] * F[1] + A[2] * F[2];
        }
        int cls = (d <= 6) ? 1 : (d <= 18) ? 2 : 3;
        fpop[d] = coe[cls] * (sum);
    }
}

void oref_node_terms(const oref_params* p, double rho, double ux, double uy, double uz,
                     double charge, double Ex, double Ey, double Ez,
                     double Fx, double Fy, double Fz, double* fe, double* he, double* fpop)
{
    node_equilibria(p, rho, ux, uy, uz, charge, Ex, Ey, Ez, fe, he);
    node_forcing(p, ux, uy, uz, Fx, Fy, Fz, fpop);
}

void oref_rates(const oref_box* b, double* out4)
{
    out4[0] = b->omega_plus; out4[1] = b->omega_minus;
    out4[2] = b->omega_c_plus; out4[3] = b->omega_c_minus;
}

void oref_set_rates(oref_box* b, const double* in4)
{
    b->omega_plus = in4[0]; b->omega_minus = in4[1];
    b->omega_c_plus = in4[2]; b->omega_c_minus = in4[3];
}

void oref_init_equilibrium(oref_box* b)
{
    /* gpu_init_equilibrium, LBM.cu:114-276 */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned z = 0; z < NZ; ++z)
        for (unsigned y = 0; y < NY; ++y)
            for (unsigned x = 0; x < NX; ++x) {
                size_t s = SIDX(x, y, z);
                double fe[27], he[27];
                node_equilibria(p, b->fld[OREF_F_RHO][s], b->fld[OREF_F_UX][s], b->fld[OREF_F_UY][s],
                                b->fld[OREF_F_UZ][s], b->fld[OREF_F_Q][s], b->fld[OREF_F_EX][s],
                                b->fld[OREF_F_EY][s], b->fld[OREF_F_EZ][s], fe, he);
                b->fld[OREF_F_F0][s] = fe[0];
                b->fld[OREF_F_H0][s] = he[0];
                for (int d = 1; d < 27; ++d) {
                    b->fld[OREF_F_F1][NIDX(x, y, z, d)] = fe[d];
                    b->fld[OREF_F_H1][NIDX(x, y, z, d)] = he[d];
                }
            }
}

/* ------------------------------------------------------------------------- */
/* LBM.                                                                       */
/* ------------------------------------------------------------------------- */

static inline double seq_sum27(const double* v)
{
    double s = v[0] + v[1];
    for (int d = 2; d < 27; ++d) s = s + v[d];
    return s;
}

static inline double mom_sum(const double* v, int a)
{
    /* (sum over c_a=+1) - (sum over c_a=-1), each in ascending order */
    double pp = v[MOM[a][0]], nn = v[MOM[a][9]];
    for (int k = 1; k < 9; ++k) { pp = pp + v[MOM[a][k]]; nn = nn + v[MOM[a][9 + k]]; }
    return pp - nn;
}

void oref_collide_save(oref_box* b)
{
    /* gpu_collide_save, LBM.cu:295-1091 */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz;
    double *f0 = b->fld[OREF_F_F0], *f1 = b->fld[OREF_F_F1], *f2 = b->fld[OREF_F_F2];
    double *h0 = b->fld[OREF_F_H0], *h1 = b->fld[OREF_F_H1], *h2 = b->fld[OREF_F_H2];
    double *ex = b->fld[OREF_F_EX], *ey = b->fld[OREF_F_EY], *ez = b->fld[OREF_F_EZ];
    double *f0bc = b->fld[OREF_F_F0BC];
    const double dt = p->dt, CFL = p->cfl, exf = p->exf;
    const double omega_plus = b->omega_plus, omega_minus = b->omega_minus;
    const double omega_c_plus = b->omega_c_plus, omega_c_minus = b->omega_c_minus;

    /* start-of-step h0 at z=1, read by the z=0 override (see file header) */
    size_t NXY = b->NXY;
    double* h0z1 = (double*)malloc(sizeof(double) * NXY);
    memcpy(h0z1, h0 + NXY, sizeof(double) * NXY);

#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned z = 0; z < NZ; ++z)
        for (unsigned y = 0; y < NY; ++y)
            for (unsigned x = 0; x < NX; ++x) {
                size_t s = SIDX(x, y, z);
                if (z == 0) f0bc[SIDX(x, y, 0)] = f0[s];          /* LBM.cu:309 */
                if (z == NZ - 1) f0bc[SIDX(x, y, 1)] = f0[s];     /* LBM.cu:311 */
                double ft[27], ht[27];
                ft[0] = f0[s]; ht[0] = h0[s];
                for (int d = 1; d < 27; ++d) { ft[d] = f1[NIDX(x, y, z, d)]; ht[d] = h1[NIDX(x, y, z, d)]; }

                double rho = seq_sum27(ft);
                double rhoinv = 1.0 / rho;
                double charge = seq_sum27(ht);
                double Ex = ex[s], Ey = ey[s], Ez = ez[s];
                double forcex = charge * Ex + exf;
                double forcey = charge * Ey;
                double forcez = charge * Ez;
                double ux = rhoinv * (mom_sum(ft, 0) / CFL + forcex * dt * 0.5);
                double uy = rhoinv * (mom_sum(ft, 1) / CFL + forcey * dt * 0.5);
                double uz = rhoinv * (mom_sum(ft, 2) / CFL + forcez * dt * 0.5);

                if (z == 0) {
                    /* wall velocity override, LBM.cu:416-493 */
                    double ftm[27], htm[27];
                    ftm[0] = 0.0;   /* rhom (LBM.cu:475) is never used */
                    htm[0] = h0z1[SIDX(x, y, 0)];
                    for (int d = 1; d < 27; ++d) { ftm[d] = f1[NIDX(x, y, 1, d)]; htm[d] = h1[NIDX(x, y, 1, d)]; }
                    double rhoinvm = 1.0 / rho;       /* LBM.cu:477 uses the wall node's rho */
                    double chargem = seq_sum27(htm);
                    size_t s1 = SIDX(x, y, 1);
                    double forcexm = chargem * ex[s1] + exf;
                    double forceym = chargem * ey[s1];
                    double forcezm = chargem * ez[s1];
                    ux = -rhoinvm * (mom_sum(ftm, 0) / CFL + forcexm * dt * 0.5);
                    uy = -rhoinvm * (mom_sum(ftm, 1) / CFL + forceym * dt * 0.5);
                    uz = -rhoinvm * (mom_sum(ftm, 2) / CFL + forcezm * dt * 0.5);
                }

                b->fld[OREF_F_RHO][s] = rho;
                b->fld[OREF_F_UX][s] = ux;
                b->fld[OREF_F_UY][s] = uy;
                b->fld[OREF_F_UZ][s] = uz;
                b->fld[OREF_F_Q][s] = charge;

                double fe[27], he[27], fpop[27];
                node_equilibria(p, rho, ux, uy, uz, charge, Ex, Ey, Ez, fe, he);
                node_forcing(p, ux, uy, uz, forcex, forcey, forcez, fpop);

                double sp = 1.0 - 0.5 * dt * omega_plus;
                double sm = 1.0 - 0.5 * dt * omega_minus;
                double tw0rp = omega_plus * dt, tw0rm = omega_minus * dt;
                double tw0cp = omega_c_plus * dt, tw0cm = omega_c_minus * dt;

                /* rest populations (LBM.cu:989, 1030-1031) */
                double fm0 = 0.0, fem0 = 0.0, hm0 = 0.0, hem0 = 0.0;
                double source0 = sp * fpop[0];
                f0[s] = ft[0] - (tw0rp * (ft[0] - fe[0]) + tw0rm * (fm0 - fem0)) + dt * source0;
                h0[s] = ht[0] - (tw0cp * (ht[0] - he[0]) + tw0cm * (hm0 - hem0));

                /* +/- pairs (d, d+1), LBM.cu:705-1089 */
                for (int d = 1; d < 27; d += 2) {
                    int e = d + 1;
                    double fp = 0.5 * (ft[d] + ft[e]);
                    double fm = 0.5 * (ft[d] - ft[e]);
                    double fep = 0.5 * (fe[d] + fe[e]);
                    double fem = 0.5 * (fe[d] - fe[e]);
                    double hp = 0.5 * (ht[d] + ht[e]);
                    double hm = 0.5 * (ht[d] - ht[e]);
                    double hep = 0.5 * (he[d] + he[e]);
                    double hem = 0.5 * (he[d] - he[e]);
                    double forcep = 0.5 * (fpop[d] + fpop[e]);
                    double forcem = 0.5 * (fpop[d] - fpop[e]);
                    double fm_e = -fm, fem_e = -fem, hm_e = -hm, hem_e = -hem, forcem_e = -forcem;
                    double source_d = sp * forcep + sm * forcem;
                    double source_e = sp * forcep + sm * forcem_e;
                    f2[NIDX(x, y, z, d)] = ft[d] - (tw0rp * (fp - fep) + tw0rm * (fm - fem)) + dt * source_d;
                    h2[NIDX(x, y, z, d)] = ht[d] - (tw0cp * (hp - hep) + tw0cm * (hm - hem));
                    f2[NIDX(x, y, z, e)] = ft[e] - (tw0rp * (fp - fep) + tw0rm * (fm_e - fem_e)) + dt * source_e;
                    h2[NIDX(x, y, z, e)] = ht[e] - (tw0cp * (hp - hep) + tw0cm * (hm_e - hem_e));
                }
            }
    free(h0z1);
}

void oref_boundary(oref_box* b)
{
    /* gpu_boundary, LBM.cu:1093-1203 */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz;
    double *f0 = b->fld[OREF_F_F0], *f1 = b->fld[OREF_F_F1], *f2 = b->fld[OREF_F_F2];
    double *h0 = b->fld[OREF_F_H0], *h2 = b->fld[OREF_F_H2], *f0bc = b->fld[OREF_F_F0BC];
    double multis = 2.0 * p->rho0 * p->uw / p->cs_square * p->ws / p->cfl;
    double multia = 2.0 * p->rho0 * p->uw / p->cs_square * p->wa / p->cfl;
    double multid = 2.0 * p->rho0 * p->uw / p->cs_square * p->wd / p->cfl;
    /* moving-wall term of the top plate per direction (LBM.cu:1145-1170):
     * +1 adds multi*, -1 subtracts it, 0 plain bounce-back */
    static const int TOPSGN[27] = {0, 1, -1, 1, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0,
                                   1, -1, 1, -1, 1, -1, -1, 1};
    for (unsigned y = 0; y < NY; ++y)
        for (unsigned x = 0; x < NX; ++x) {
            /* lower plate, LBM.cu:1103-1132 */
            f0[SIDX(x, y, 0)] = f0bc[SIDX(x, y, 0)];
            for (int d = 1; d < 27; ++d) f2[NIDX(x, y, 0, d)] = f1[NIDX(x, y, 0, opp_of(d))];
            /* upper plate, LBM.cu:1142-1201 */
            f0[SIDX(x, y, NZ - 1)] = f0bc[SIDX(x, y, 1)];
            for (int d = 1; d < 27; ++d) {
                double m = (d <= 6) ? multis : (d <= 18) ? multia : multid;
                double v = f1[NIDX(x, y, NZ - 1, opp_of(d))];
                if (TOPSGN[d] > 0) v = v + m;
                else if (TOPSGN[d] < 0) v = v - m;
                f2[NIDX(x, y, NZ - 1, d)] = v;
            }
            h0[SIDX(x, y, NZ - 1)] = h0[SIDX(x, y, NZ - 2)];
            for (int d = 1; d < 27; ++d) h2[NIDX(x, y, NZ - 1, d)] = h2[NIDX(x, y, NZ - 2, d)];
        }
}

void oref_stream(oref_box* b)
{
    /* gpu_stream, LBM.cu:1205-1279: periodic pull in x, y and z */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz;
    double *f1 = b->fld[OREF_F_F1], *f2 = b->fld[OREF_F_F2];
    double *h1 = b->fld[OREF_F_H1], *h2 = b->fld[OREF_F_H2];
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned z = 0; z < NZ; ++z)
        for (unsigned y = 0; y < NY; ++y)
            for (unsigned x = 0; x < NX; ++x)
                for (int d = 1; d < 27; ++d) {
                    unsigned xs = (NX + x - CV[d][0]) % NX;
                    unsigned ys = (NY + y - CV[d][1]) % NY;
                    unsigned zs = (NZ + z - CV[d][2]) % NZ;
                    f1[NIDX(x, y, z, d)] = f2[NIDX(xs, ys, zs, d)];
                    h1[NIDX(x, y, z, d)] = h2[NIDX(xs, ys, zs, d)];
                }
}

void oref_bc_charge(oref_box* b)
{
    /* gpu_bc_charge, LBM.cu:1281-1353: anti-bounce-back, q = charge0 at z=0 */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz;
    (void)NZ;
    double *h0 = b->fld[OREF_F_H0], *h1 = b->fld[OREF_F_H1], *h2 = b->fld[OREF_F_H2];
    double multi0c = 2.0 * p->charge0 * p->w0;
    double multisc = 2.0 * p->charge0 * p->ws;
    double multiac = 2.0 * p->charge0 * p->wa;
    double multidc = 2.0 * p->charge0 * p->wd;
    for (unsigned y = 0; y < NY; ++y)
        for (unsigned x = 0; x < NX; ++x) {
            double ht[27];
            for (int d = 1; d < 27; ++d) ht[d] = h2[NIDX(x, y, 0, d)];
            h0[SIDX(x, y, 0)] = -h0[SIDX(x, y, 0)] + multi0c;
            for (int d = 1; d < 27; ++d) {
                double m = (d <= 6) ? multisc : (d <= 18) ? multiac : multidc;
                h1[NIDX(x, y, 0, d)] = -ht[opp_of(d)] + m;
            }
        }
}

void oref_stream_collide_save(oref_box* b)
{
    /* LBM.cu:278-293 */
    oref_collide_save(b);
    oref_boundary(b);
    oref_stream(b);
    oref_bc_charge(b);
}

/* ------------------------------------------------------------------------- */
/* Poisson.                                                                   */
/* ------------------------------------------------------------------------- */

void oref_extension(oref_box* b)
{
    /* odd_extension, poisson.cu:87-121 */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz, NE = p->ne;
    double* c = b->fld[OREF_F_Q];
    double* ce = b->fld[OREF_F_SEXT];
    double eps = p->eps, voltage = p->voltage, dz = p->dz;
    for (unsigned z = 0; z < NE; ++z)
        for (unsigned y = 0; y < NY; ++y)
            for (unsigned x = 0; x < NX; ++x) {
                size_t i = SIDX(x, y, z);
                double re;
                if (z == 0) re = 0.0;
                else if (z == 1) re = -c[SIDX(x, y, z)] / eps - voltage / dz / dz;
                else if (z > 1 && z < NZ - 1) re = -c[SIDX(x, y, z)] / eps;
                else if (z == NZ - 1) re = 0.0;
                else if (z > NZ - 1 && z < NE - 1) re = c[SIDX(x, y, NE - z)] / eps;
                else re = c[SIDX(x, y, 1)] / eps + voltage / dz / dz;
                ce[2 * i] = re;
                ce[2 * i + 1] = 0.0;
            }
}

void oref_fft3d(oref_box* b, double* in, double* out, int sign)
{
    const oref_params* p = &b->p;
    if (out != in) memcpy(out, in, sizeof(double) * 2 * b->SE);
    fft_axis(b->px, out, p->nx, p->ny, p->ne, 0, sign);
    fft_axis(b->py, out, p->nx, p->ny, p->ne, 1, sign);
    fft_axis(b->pz, out, p->nx, p->ny, p->ne, 2, sign);
}

void oref_derivative(oref_box* b)
{
    /* gpu_derivative, poisson.cu:132-143 */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NE = p->ne;
    double* src = b->fld[OREF_F_FREQ];
    double *kx = b->fld[OREF_F_KX], *ky = b->fld[OREF_F_KY], *kz = b->fld[OREF_F_KZ];
    double dz = p->dz;
#pragma omp parallel for schedule(static)
    for (unsigned z = 0; z < NE; ++z)
        for (unsigned y = 0; y < NY; ++y)
            for (unsigned x = 0; x < NX; ++x) {
                double I = kx[x], J = ky[y], K = kz[z];
                double mu = (4.0 / dz / dz) * (sin(K * dz * 0.5) * sin(K * dz * 0.5)) + I * I + J * J;
                if (y == 0 && x == 0 && z == 0) mu = 1.0;
                size_t i = SIDX(x, y, z);
                src[2 * i] = -src[2 * i] / mu;
                src[2 * i + 1] = -src[2 * i + 1] / mu;
            }
}

void oref_extract(oref_box* b)
{
    /* odd_extract, poisson.cu:154-167 */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz;
    double* phi = b->fld[OREF_F_PHI];
    double* pe = b->fld[OREF_F_PHIEXT];
    double size = (double)((size_t)p->nx * p->ny * p->ne);
    for (unsigned z = 0; z < NZ; ++z)
        for (unsigned y = 0; y < NY; ++y)
            for (unsigned x = 0; x < NX; ++x) {
                size_t i = SIDX(x, y, z);
                if (z == 0) phi[i] = p->voltage;
                else if (z == NZ - 1) phi[i] = 0.0;
                else phi[i] = pe[2 * i] / size;
            }
}

void oref_efield(oref_box* b)
{
    /* gpu_efield + gpu_bc, poisson.cu:44-73 */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz;
    double* fi = b->fld[OREF_F_PHI];
    double *ex = b->fld[OREF_F_EX], *ey = b->fld[OREF_F_EY], *ez = b->fld[OREF_F_EZ];
    double dx = p->dx, dy = p->dy, dz = p->dz;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned z = 0; z < NZ; ++z)
        for (unsigned y = 0; y < NY; ++y)
            for (unsigned x = 0; x < NX; ++x) {
                unsigned xp1 = (x + 1) % NX, yp1 = (y + 1) % NY, zp1 = (z + 1) % NZ;
                unsigned xm1 = (NX + x - 1) % NX, ym1 = (NY + y - 1) % NY, zm1 = (NZ + z - 1) % NZ;
                ex[SIDX(x, y, z)] = 0.5 * (fi[SIDX(xm1, y, z)] - fi[SIDX(xp1, y, z)]) / dx;
                ey[SIDX(x, y, z)] = 0.5 * (fi[SIDX(x, ym1, z)] - fi[SIDX(x, yp1, z)]) / dy;
                ez[SIDX(x, y, z)] = 0.5 * (fi[SIDX(x, y, zm1)] - fi[SIDX(x, y, zp1)]) / dz;
            }
    for (unsigned y = 0; y < NY; ++y)
        for (unsigned x = 0; x < NX; ++x) {
            ez[SIDX(x, y, 0)] = ez[SIDX(x, y, 1)];
            ez[SIDX(x, y, NZ - 1)] = ez[SIDX(x, y, NZ - 2)];
        }
}

void oref_poisson(oref_box* b)
{
    /* main.cu:191-203 */
    oref_extension(b);
    oref_fft3d(b, b->fld[OREF_F_SEXT], b->fld[OREF_F_FREQ], -1);
    oref_derivative(b);
    oref_fft3d(b, b->fld[OREF_F_FREQ], b->fld[OREF_F_PHIEXT], +1);
    oref_extract(b);
}

void oref_poisson_tridiag(oref_box* b)
{
    /* Same discrete problem as oref_poisson: spectral Laplacian in x,y
     * (k = kx[i], ky[j] of main.cu:109-125), second-order differences in z,
     * phi(0)=voltage, phi(nz-1)=0.  phi = voltage*(1 - z/(nz-1)) + phi_q. */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz;
    size_t NXY = b->NXY;
    int n = (int)NZ - 2;
    double* a = (double*)malloc(sizeof(double) * 2 * NXY * (size_t)n);
    double* q = b->fld[OREF_F_Q];
    for (int k = 0; k < n; ++k)
        for (size_t i = 0; i < NXY; ++i) {
            a[2 * (k * NXY + i)] = -q[(size_t)(k + 1) * NXY + i] / p->eps;
            a[2 * (k * NXY + i) + 1] = 0.0;
        }
    fft_axis(b->px, a, NX, NY, n, 0, -1);
    fft_axis(b->py, a, NX, NY, n, 1, -1);
    double dz2 = p->dz * p->dz;
    double *kx = b->fld[OREF_F_KX], *ky = b->fld[OREF_F_KY];
#pragma omp parallel
    {
        double* cp = (double*)malloc(sizeof(double) * (size_t)n);
        double* dr = (double*)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for schedule(static)
        for (long long ii = 0; ii < (long long)NXY; ++ii) {
            size_t i = (size_t)ii;
            double k2 = kx[i % NX] * kx[i % NX] + ky[i / NX] * ky[i / NX];
            double bdiag = -(2.0 + k2 * dz2);
            for (int k = 0; k < n; ++k) {
                double rr = dz2 * a[2 * (k * NXY + i)], ri = dz2 * a[2 * (k * NXY + i) + 1];
                if (k == 0) {
                    cp[0] = 1.0 / bdiag;
                    dr[0] = rr * cp[0]; dr[1] = ri * cp[0];
                } else {
                    double den = 1.0 / (bdiag - cp[k - 1]);
                    cp[k] = den;
                    dr[2 * k] = (rr - dr[2 * (k - 1)]) * den;
                    dr[2 * k + 1] = (ri - dr[2 * (k - 1) + 1]) * den;
                }
            }
            for (int k = n - 1; k >= 0; --k) {
                if (k < n - 1) {
                    dr[2 * k] -= cp[k] * dr[2 * (k + 1)];
                    dr[2 * k + 1] -= cp[k] * dr[2 * (k + 1) + 1];
                }
                a[2 * (k * NXY + i)] = dr[2 * k];
                a[2 * (k * NXY + i) + 1] = dr[2 * k + 1];
            }
        }
        free(cp); free(dr);
    }
    fft_axis(b->px, a, NX, NY, n, 0, +1);
    fft_axis(b->py, a, NX, NY, n, 1, +1);
    double* phi = b->fld[OREF_F_PHI];
    double inv = 1.0 / (double)NXY;
    for (unsigned z = 0; z < NZ; ++z) {
        double lin = p->voltage * (1.0 - (double)z / (double)(NZ - 1));
        for (size_t i = 0; i < NXY; ++i) {
            if (z == 0) phi[i] = p->voltage;
            else if (z == NZ - 1) phi[(size_t)z * NXY + i] = 0.0;
            else phi[(size_t)z * NXY + i] = lin + a[2 * ((z - 1) * NXY + i)] * inv;
        }
    }
    free(a);
}

/* ------------------------------------------------------------------------- */
/* Coupled step and diagnostics.                                              */
/* ------------------------------------------------------------------------- */

void oref_step(oref_box* b, int nsteps)
{
    for (int i = 0; i < nsteps; ++i) {
        oref_stream_collide_save(b);   /* main.cu:184 */
        oref_poisson(b);               /* main.cu:191-203 */
        oref_efield(b);                /* main.cu:206 */
    }
}

void oref_step_tridiag(oref_box* b, int nsteps)
{
    for (int i = 0; i < nsteps; ++i) {
        oref_stream_collide_save(b);
        oref_poisson_tridiag(b);
        oref_efield(b);
    }
}

double oref_current(oref_box* b)
{
    /* LBM.cu:1646-1671, applied to copies so the box state is unchanged */
    const oref_params* p = &b->p;
    unsigned NX = p->nx, NY = p->ny, NZ = p->nz;
    const double* c = b->fld[OREF_F_Q];
    const double* ez = b->fld[OREF_F_EZ];
    double I = 0;
    for (unsigned y = 0; y < NY; y++)
        for (unsigned x = 0; x < NX; x++) {
            double ctop = 2.0 * c[SIDX(x, y, NZ - 2)] - c[SIDX(x, y, NZ - 3)];
            I += ctop * ez[SIDX(x, y, NZ - 1)];
        }
    I = I * p->K * p->dy * p->dy;
    return I;
}
