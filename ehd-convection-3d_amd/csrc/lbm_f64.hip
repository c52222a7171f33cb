// lbm_f64.hip -- fused pull-scheme LBM step in FP64, in the reference's
// operation order, plus the population init/read/write kernels of both
// storage precisions.
//
// Compiled with -ffp-contract=off: every product and sum is rounded on its
// own, in the order written here, which is the order of gpu_collide_save
// (LBM.cu:295-1091).  With the same inputs the FP64 step is bit-identical to
// the reference's collide -> boundary -> stream -> bc_charge sequence
// (LBM.cu:278-293).
//
// One launch does the whole LBM part of a step for every node:
//   * pull: the pre-collision population d of node x is the post-collision,
//     post-boundary value of node x - c_d (periodic in x, y, z; the z wrap
//     connects the two wall planes exactly as gpu_stream does);
//   * z=0: h is rebuilt by anti-bounce-back from the node's own
//     post-collision h (gpu_bc_charge), the velocity is overridden from the
//     pre-collision populations of z=1 (LBM.cu:416-493);
//   * collide (TRT, Guo force for f, drift u+K*E for h);
//   * walls: f is stored bounced-back (gpu_boundary), the rest population is
//     frozen, the z=nz-2 node also stores its h into the z=nz-1 slots (the
//     zero-gradient copy of gpu_boundary).
// The state between steps is therefore the post-collision, post-boundary f
// and the post-collision h; a double buffer (A read, B written) replaces the
// reference's f1/f2 pair and the in-place rest populations.  After an
// initialisation or an explicit write the state holds pre-collision values
// ("direct" representation) and the first launch reads them in place.
#include <hip/hip_runtime.h>

#include "ehd_internal.h"

namespace ehd {
namespace {

__device__ __forceinline__ double seq27(const double* v)
{
    double s = v[0] + v[1];
#pragma unroll
    for (int d = 2; d < 27; ++d) s = s + v[d];
    return s;
}

__device__ __forceinline__ double momsum(const double* v, int a)
{
    double pp = v[MOM[a][0]], nn = v[MOM[a][9]];
#pragma unroll
    for (int k = 1; k < 9; ++k) {
        pp = pp + v[MOM[a][k]];
        nn = nn + v[MOM[a][9 + k]];
    }
    return pp - nn;
}

template <bool H>
__device__ __forceinline__ double cdot(int d, const double* t)
{
    const int* tm = H ? EQH[d] : EQF[d];
    int k = tm[0];
    double acc = (k > 0) ? t[k] : -t[-k];
    if (tm[1] != 0) { k = tm[1]; acc = (k > 0) ? acc + t[k] : acc - t[-k]; }
    if (tm[2] != 0) { k = tm[2]; acc = (k > 0) ? acc + t[k] : acc - t[-k]; }
    return acc;
}

// Guo force population of direction d (LBM.cu:671-702), same association.
__device__ __forceinline__ double fpop_dir(int d, const double* u, const double* F, double cflinv,
                                           double cflinv2, double coe)
{
    int ax[3] = {0, 0, 0}, zr[3] = {0, 0, 0};
    int na = 0, nzr = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (cv(d, a) != 0) ax[na++] = a;
        else zr[nzr++] = a;
    }
    double A[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (i < na) {
            int a = ax[i];
            double P = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                if (j < na) {
                    int b = ax[j];
                    int s = cv(d, a) * cv(d, b);
                    if (j == 0) P = (s > 0) ? u[b] : -u[b];
                    else P = (s > 0) ? P + u[b] : P - u[b];
                }
            }
            double head = (cv(d, a) > 0) ? (cflinv - u[a]) : (-cflinv - u[a]);
            A[i] = head + P * cflinv2;
        }
    }
    double sum;
    if (na == 1) sum = -u[zr[0]] * F[zr[0]] - u[zr[1]] * F[zr[1]] + A[0] * F[ax[0]];
    else if (na == 2) sum = A[0] * F[ax[0]] + A[1] * F[ax[1]] - u[zr[0]] * F[zr[0]];
    else sum = A[0] * F[0] + A[1] * F[1] + A[2] * F[2];
    return coe * (sum);
}

// TRT relaxation of one node (LBM.cu:519-1089): equilibria, force
// populations, +/- splits, sources, relaxation.
__device__ __forceinline__ void relax_f64(const LbmParams& P, const double* ft, const double* ht,
                                          double rho, double charge, double ux, double uy, double uz,
                                          double Ex, double Ey, double Ez, double forcex, double forcey,
                                          double forcez, double* fo, double* ho)
{
    const double cs_square = P.cs_square, CFL = P.cfl, K = P.K, dt = P.dt;
    const double wr[4] = {P.w[0] * rho, P.w[1] * rho, P.w[2] * rho, P.w[3] * rho};
    const double wc[4] = {P.w[0] * charge, P.w[1] * charge, P.w[2] * charge, P.w[3] * charge};
    const double omusq = 1.0 - 0.5 * (ux * ux + uy * uy + uz * uz) / cs_square;
    const double omusq_c = 1.0 - 0.5 * ((ux + K * Ex) * (ux + K * Ex) + (uy + K * Ey) * (uy + K * Ey)
                                        + (uz + K * Ez) * (uz + K * Ez)) / cs_square;
    double t[4], tc[4];
    t[0] = 0.0; tc[0] = 0.0;
    t[1] = ux / cs_square / CFL;
    t[2] = uy / cs_square / CFL;
    t[3] = uz / cs_square / CFL;
    tc[1] = (ux + K * Ex) / cs_square / CFL;
    tc[2] = (uy + K * Ey) / cs_square / CFL;
    tc[3] = (uz + K * Ez) / cs_square / CFL;

    const double coe[4] = {P.w[0] / cs_square, P.w[1] / cs_square, P.w[2] / cs_square, P.w[3] / cs_square};
    const double cflinv = 1.0 / CFL;
    const double cflinv2 = cflinv * cflinv / cs_square;
    const double u[3] = {ux, uy, uz};
    const double F[3] = {forcex, forcey, forcez};

    const double sp = 1.0 - 0.5 * dt * P.omega_plus;
    const double sm = 1.0 - 0.5 * dt * P.omega_minus;
    const double tw0rp = P.omega_plus * dt, tw0rm = P.omega_minus * dt;
    const double tw0cp = P.omega_c_plus * dt, tw0cm = P.omega_c_minus * dt;

    // rest populations
    {
        const double fe0 = wr[0] * (omusq);
        const double he0 = wc[0] * (omusq_c);
        const double fpop0 = -coe[0] * (ux * forcex + uy * forcey + uz * forcez);
        const double zero = 0.0;
        const double source0 = sp * fpop0;
        fo[0] = ft[0] - (tw0rp * (ft[0] - fe0) + tw0rm * (zero - zero)) + dt * source0;
        ho[0] = ht[0] - (tw0cp * (ht[0] - he0) + tw0cm * (zero - zero));
    }
#pragma unroll
    for (int d = 1; d < 27; d += 2) {
        const int e = d + 1;
        const int c = WCLS[d];
        const double cud = cdot<false>(d, t), cue = cdot<false>(e, t);
        const double fed = wr[c] * (omusq + cud * (1.0 + 0.5 * cud));
        const double fee = wr[c] * (omusq + cue * (1.0 + 0.5 * cue));
        const double ccd = cdot<true>(d, tc), cce = cdot<true>(e, tc);
        const double hed = wc[c] * (omusq_c + ccd * (1.0 + 0.5 * ccd));
        const double hee = wc[c] * (omusq_c + cce * (1.0 + 0.5 * cce));
        const double fpd = fpop_dir(d, u, F, cflinv, cflinv2, coe[c]);
        const double fpe = fpop_dir(e, u, F, cflinv, cflinv2, coe[c]);

        const double fp = 0.5 * (ft[d] + ft[e]);
        const double fm = 0.5 * (ft[d] - ft[e]);
        const double fep = 0.5 * (fed + fee);
        const double fem = 0.5 * (fed - fee);
        const double hp = 0.5 * (ht[d] + ht[e]);
        const double hm = 0.5 * (ht[d] - ht[e]);
        const double hep = 0.5 * (hed + hee);
        const double hem = 0.5 * (hed - hee);
        const double forcep = 0.5 * (fpd + fpe);
        const double forcem = 0.5 * (fpd - fpe);
        const double fm_e = -fm, fem_e = -fem, hm_e = -hm, hem_e = -hem, forcem_e = -forcem;
        const double source_d = sp * forcep + sm * forcem;
        const double source_e = sp * forcep + sm * forcem_e;
        fo[d] = ft[d] - (tw0rp * (fp - fep) + tw0rm * (fm - fem)) + dt * source_d;
        ho[d] = ht[d] - (tw0cp * (hp - hep) + tw0cm * (hm - hem));
        fo[e] = ft[e] - (tw0rp * (fp - fep) + tw0rm * (fm_e - fem_e)) + dt * source_e;
        ho[e] = ht[e] - (tw0cp * (hp - hep) + tw0cm * (hm_e - hem_e));
    }
}

// Two waves per SIMD: left to itself the compiler gives the steady-state
// variant 256 VGPRs + 4 AGPRs, one wave per SIMD; capped it keeps 256 VGPRs
// and spills 5 dwords per lane to scratch.
template <bool FIRST, bool EBUF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
lbm_f64_kernel(LbmParams P, PopBufs pb, MacroBufs mb, int macro)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.S) return;
    const unsigned nx = P.nx, ny = P.ny, nz = P.nz, nxy = P.nxy;
    const unsigned x = s % nx, t = s / nx, y = t % ny, z = t / ny;
    const unsigned xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
    const unsigned ym = (y == 0) ? ny - 1 : y - 1, yp = (y == ny - 1) ? 0 : y + 1;
    const unsigned zm = (z == 0) ? nz - 1 : z - 1, zp = (z == nz - 1) ? 0 : z + 1;
    const unsigned XS[3] = {xm, x, xp};
    // populations in row blocks [z][y][d-1][x]; a pull along c reads row (y - c_y, z - c_z)
    const unsigned RW = 26u * nx;
    const unsigned YR[3] = {ym, y, yp}, ZR[3] = {zm * ny, z * ny, zp * ny};
    const unsigned own = t * RW + x;

    const double* __restrict__ fA = (const double*)pb.f;
    const double* __restrict__ f0A = (const double*)pb.f0;
    const double* __restrict__ hA = (const double*)pb.h;
    const double* __restrict__ h0A = (const double*)pb.h0;

    double ft[27], ht[27];
    ft[0] = f0A[s];
    if (FIRST) {
        ht[0] = h0A[s];
#pragma unroll
        for (int d = 1; d < 27; ++d) {
            ft[d] = fA[own + (d - 1) * nx];
            ht[d] = hA[own + (d - 1) * nx];
        }
    } else {
#pragma unroll
        for (int d = 1; d < 27; ++d)
            ft[d] = fA[(ZR[1 - CZ[d]] + YR[1 - CY[d]]) * RW + (d - 1) * nx + XS[1 - CX[d]]];
        if (z == 0) {
            ht[0] = -h0A[s] + P.inj[0];
#pragma unroll
            for (int d = 1; d < 27; ++d) ht[d] = -hA[own + (opp(d) - 1) * nx] + P.inj[WCLS[d]];
        } else {
            ht[0] = h0A[s];
#pragma unroll
            for (int d = 1; d < 27; ++d)
                ht[d] = hA[(ZR[1 - CZ[d]] + YR[1 - CY[d]]) * RW + (d - 1) * nx + XS[1 - CX[d]]];
        }
    }

    // E = -grad(phi) of the previous potential (gpu_efield + gpu_bc)
    const double* phi = (const double*)mb.phi;
    double Ex, Ey, Ez;
    if (EBUF) {
        Ex = mb.ex[s]; Ey = mb.ey[s]; Ez = mb.ez[s];
    } else {
        const unsigned yc = y * nx, zc = z * nxy;
        Ex = 0.5 * (phi[xm + yc + zc] - phi[xp + yc + zc]) / P.dx;
        Ey = 0.5 * (phi[x + ym * nx + zc] - phi[x + yp * nx + zc]) / P.dy;
        const unsigned ze = (z == 0) ? 1u : ((z == nz - 1) ? nz - 2 : z);
        Ez = 0.5 * (phi[x + yc + (ze - 1) * nxy] - phi[x + yc + (ze + 1) * nxy]) / P.dz;
    }

    const double dt = P.dt, CFL = P.cfl;
    const double rho = seq27(ft);
    const double rhoinv = 1.0 / rho;
    const double charge = seq27(ht);
    const double forcex = charge * Ex + P.exf;
    const double forcey = charge * Ey;
    const double forcez = charge * Ez;
    double ux = rhoinv * (momsum(ft, 0) / CFL + forcex * dt * 0.5);
    double uy = rhoinv * (momsum(ft, 1) / CFL + forcey * dt * 0.5);
    double uz = rhoinv * (momsum(ft, 2) / CFL + forcez * dt * 0.5);

    if (z == 0) {
        // velocity override from the pre-collision populations of z=1
        double ftm[27], htm[27];
        const unsigned s1 = s + nxy;
        ftm[0] = 0.0;
        htm[0] = h0A[s1];
        if (FIRST) {
            const unsigned own1 = own + ny * RW;
#pragma unroll
            for (int d = 1; d < 27; ++d) {
                ftm[d] = fA[own1 + (d - 1) * nx];
                htm[d] = hA[own1 + (d - 1) * nx];
            }
        } else {
#pragma unroll
            for (int d = 1; d < 27; ++d) {
                // node (x, y, 1) pulls from row (y - c_y, 1 - c_z)
                const unsigned src = ((unsigned)(1 - CZ[d]) * ny + YR[1 - CY[d]]) * RW + (d - 1) * nx + XS[1 - CX[d]];
                ftm[d] = fA[src];
                htm[d] = hA[src];
            }
        }
        double Exm, Eym, Ezm;
        if (EBUF) {
            Exm = mb.ex[s1]; Eym = mb.ey[s1]; Ezm = mb.ez[s1];
        } else {
            const unsigned yc = y * nx;
            Exm = 0.5 * (phi[xm + yc + nxy] - phi[xp + yc + nxy]) / P.dx;
            Eym = 0.5 * (phi[x + ym * nx + nxy] - phi[x + yp * nx + nxy]) / P.dy;
            Ezm = 0.5 * (phi[x + yc] - phi[x + yc + 2 * nxy]) / P.dz;
        }
        const double rhoinvm = 1.0 / rho;   // the wall node's rho (LBM.cu:477)
        const double chargem = seq27(htm);
        const double forcexm = chargem * Exm + P.exf;
        const double forceym = chargem * Eym;
        const double forcezm = chargem * Ezm;
        ux = -rhoinvm * (momsum(ftm, 0) / CFL + forcexm * dt * 0.5);
        uy = -rhoinvm * (momsum(ftm, 1) / CFL + forceym * dt * 0.5);
        uz = -rhoinvm * (momsum(ftm, 2) / CFL + forcezm * dt * 0.5);
    }

    if (macro) {
        ((double*)mb.rho)[s] = rho;
        ((double*)mb.ux)[s] = ux;
        ((double*)mb.uy)[s] = uy;
        ((double*)mb.uz)[s] = uz;
    }
    ((double*)mb.q)[s] = charge;

    double fo[27], ho[27];
    relax_f64(P, ft, ht, rho, charge, ux, uy, uz, Ex, Ey, Ez, forcex, forcey, forcez, fo, ho);

    double* __restrict__ fB = (double*)pb.fo;
    double* __restrict__ f0B = (double*)pb.f0o;
    double* __restrict__ hB = (double*)pb.ho;
    double* __restrict__ h0B = (double*)pb.h0o;
    if (z == 0) {
        f0B[s] = ft[0];
#pragma unroll
        for (int d = 1; d < 27; ++d) fB[own + (d - 1) * nx] = ft[opp(d)];
    } else if (z == nz - 1) {
        f0B[s] = ft[0];
#pragma unroll
        for (int d = 1; d < 27; ++d) {
            double v = ft[opp(d)];
            if (TOPSGN[d] > 0) v = v + P.multi[WCLS[d]];
            else if (TOPSGN[d] < 0) v = v - P.multi[WCLS[d]];
            fB[own + (d - 1) * nx] = v;
        }
    } else {
        f0B[s] = fo[0];
#pragma unroll
        for (int d = 1; d < 27; ++d) fB[own + (d - 1) * nx] = fo[d];
    }
    if (z != nz - 1) {
        h0B[s] = ho[0];
#pragma unroll
        for (int d = 1; d < 27; ++d) hB[own + (d - 1) * nx] = ho[d];
        if (z == nz - 2) {
            h0B[s + nxy] = ho[0];
#pragma unroll
            for (int d = 1; d < 27; ++d) hB[own + ny * RW + (d - 1) * nx] = ho[d];
        }
    }
}

// Index of moving population d (1..26) of node `node` in the box's row-block
// layout [z][y][d-1][x].
__device__ __forceinline__ size_t pop_at(int d, unsigned node, unsigned nx)
{
    return (size_t)(node / nx) * 26u * nx + (size_t)(d - 1) * nx + node % nx;
}

// Equilibrium populations (gpu_init_equilibrium, LBM.cu:114-276) in FP64,
// stored as doubles (F64) or as float deviations from w*rs[z], w*qs[z] (F32S).
template <typename T>
__global__ void __launch_bounds__(256) init_eq_kernel(LbmParams P, const double* rhoA, const double* qA,
                                                      const double* uxA, const double* uyA,
                                                      const double* uzA, const double* exA,
                                                      const double* eyA, const double* ezA, T* f, T* f0,
                                                      T* h, T* h0)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.S) return;
    const unsigned z = s / P.nxy;
    const double cs_square = P.cs_square, CFL = P.cfl, K = P.K;
    const double rho = rhoA[s], ux = uxA[s], uy = uyA[s], uz = uzA[s], charge = qA[s];
    const double Ex = exA[s], Ey = eyA[s], Ez = ezA[s];
    const double wr[4] = {P.w[0] * rho, P.w[1] * rho, P.w[2] * rho, P.w[3] * rho};
    const double wc[4] = {P.w[0] * charge, P.w[1] * charge, P.w[2] * charge, P.w[3] * charge};
    const double omusq = 1.0 - 0.5 * (ux * ux + uy * uy + uz * uz) / cs_square;
    const double omusq_c = 1.0 - 0.5 * ((ux + K * Ex) * (ux + K * Ex) + (uy + K * Ey) * (uy + K * Ey)
                                        + (uz + K * Ez) * (uz + K * Ez)) / cs_square;
    double t[4], tc[4];
    t[0] = 0.0; tc[0] = 0.0;
    t[1] = ux / cs_square / CFL;
    t[2] = uy / cs_square / CFL;
    t[3] = uz / cs_square / CFL;
    tc[1] = (ux + K * Ex) / cs_square / CFL;
    tc[2] = (uy + K * Ey) / cs_square / CFL;
    tc[3] = (uz + K * Ez) / cs_square / CFL;
    const double rs = P.rs ? P.rs[z] : 0.0, qs = P.qs ? P.qs[z] : 0.0;
    f0[s] = (T)(wr[0] * (omusq) - P.w[0] * rs);
    h0[s] = (T)(wc[0] * (omusq_c) - P.w[0] * qs);
#pragma unroll
    for (int d = 1; d < 27; ++d) {
        const int c = WCLS[d];
        const double cu = cdot<false>(d, t);
        const double cc = cdot<true>(d, tc);
        const double fe = wr[c] * (omusq + cu * (1.0 + 0.5 * cu));
        const double he = wc[c] * (omusq_c + cc * (1.0 + 0.5 * cc));
        f[pop_at(d, s, P.nx)] = (T)(fe - P.w[c] * rs);
        h[pop_at(d, s, P.nx)] = (T)(he - P.w[c] * qs);
    }
}

// Reference pre-collision state (f0, f1, h0, h1 after stream and bc_charge)
// as doubles, from the box's representation.  A stored value of plane z
// carries the shift w*rs[z] (w*qs[z]); F64 has no shift.
template <typename T>
__global__ void __launch_bounds__(256) read_pre_kernel(LbmParams P, int direct, const T* f, const T* f0,
                                                       const T* h, const T* h0, double* f1o, double* f0o,
                                                       double* h1o, double* h0o)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.S) return;
    const unsigned nx = P.nx, ny = P.ny, nz = P.nz, nxy = P.nxy, S = P.S;
    const unsigned x = s % nx, t = s / nx, y = t % ny, z = t / ny;
    const unsigned xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
    const unsigned ym = (y == 0) ? ny - 1 : y - 1, yp = (y == ny - 1) ? 0 : y + 1;
    const unsigned zm = (z == 0) ? nz - 1 : z - 1, zp = (z == nz - 1) ? 0 : z + 1;
    const unsigned XS[3] = {xm, x, xp};
    const unsigned YS[3] = {ym * nx, y * nx, yp * nx};
    const unsigned ZS[3] = {zm * nxy, z * nxy, zp * nxy};
    const unsigned ZP[3] = {zm, z, zp};
    auto rsh = [&](unsigned zz) { return P.rs ? P.rs[zz] : 0.0; };
    auto qsh = [&](unsigned zz) { return P.qs ? P.qs[zz] : 0.0; };
    f0o[s] = (double)f0[s] + P.w[0] * rsh(z);
    if (!direct && z == 0) h0o[s] = -((double)h0[s] + P.w[0] * qsh(0)) + P.inj[0];
    else h0o[s] = (double)h0[s] + P.w[0] * qsh(z);
#pragma unroll
    for (int d = 1; d < 27; ++d) {
        const double wf = P.w[WCLS[d]];
        const unsigned src = direct ? s : (XS[1 - CX[d]] + YS[1 - CY[d]] + ZS[1 - CZ[d]]);
        const unsigned zs = direct ? z : ZP[1 - CZ[d]];
        f1o[(d - 1) * S + s] = (double)f[pop_at(d, src, nx)] + wf * rsh(zs);
        if (!direct && z == 0)
            h1o[(d - 1) * S + s] = -((double)h[pop_at(opp(d), s, nx)] + wf * qsh(0)) + P.inj[WCLS[d]];
        else
            h1o[(d - 1) * S + s] = (double)h[pop_at(d, src, nx)] + wf * qsh(zs);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) write_pre_kernel(LbmParams P, const double* f1, const double* f0,
                                                        const double* h1, const double* h0, T* f, T* f0o, T* h,
                                                        T* h0o)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.S) return;
    const unsigned S = P.S;
    const unsigned z = s / P.nxy;
    const double rs = P.rs ? P.rs[z] : 0.0, qs = P.qs ? P.qs[z] : 0.0;
    f0o[s] = (T)(f0[s] - P.w[0] * rs);
    h0o[s] = (T)(h0[s] - P.w[0] * qs);
#pragma unroll
    for (int d = 1; d < 27; ++d) {
        const double wf = P.w[WCLS[d]];
        f[pop_at(d, s, P.nx)] = (T)(f1[(d - 1) * S + s] - wf * rs);
        h[pop_at(d, s, P.nx)] = (T)(h1[(d - 1) * S + s] - wf * qs);
    }
}

inline dim3 grid_for(unsigned n) { return dim3((n + 255) / 256); }

}  // namespace

hipError_t launch_lbm_f64(const LbmParams& P, const PopBufs& pb, const MacroBufs& mb, bool first,
                          bool ebuf, bool macro, hipStream_t st)
{
    dim3 g = grid_for(P.S), b(256);
    int m = macro ? 1 : 0;
    if (first && ebuf) lbm_f64_kernel<true, true><<<g, b, 0, st>>>(P, pb, mb, m);
    else if (first) lbm_f64_kernel<true, false><<<g, b, 0, st>>>(P, pb, mb, m);
    else if (ebuf) lbm_f64_kernel<false, true><<<g, b, 0, st>>>(P, pb, mb, m);
    else lbm_f64_kernel<false, false><<<g, b, 0, st>>>(P, pb, mb, m);
    return hipGetLastError();
}

hipError_t launch_init_equilibrium(Prec prec, const LbmParams& P, const double* rho, const double* q,
                                   const double* ux, const double* uy, const double* uz,
                                   const double* ex, const double* ey, const double* ez, void* f,
                                   void* f0, void* h, void* h0, hipStream_t st)
{
    dim3 g = grid_for(P.S), b(256);
    if (prec == Prec::F64)
        init_eq_kernel<double><<<g, b, 0, st>>>(P, rho, q, ux, uy, uz, ex, ey, ez, (double*)f, (double*)f0,
                                                (double*)h, (double*)h0);
    else
        init_eq_kernel<float><<<g, b, 0, st>>>(P, rho, q, ux, uy, uz, ex, ey, ez, (float*)f, (float*)f0,
                                               (float*)h, (float*)h0);
    return hipGetLastError();
}

hipError_t launch_read_pre(Prec prec, const LbmParams& P, bool direct, const void* f, const void* f0,
                           const void* h, const void* h0, double* f1o, double* f0o, double* h1o,
                           double* h0o, hipStream_t st)
{
    dim3 g = grid_for(P.S), b(256);
    if (prec == Prec::F64)
        read_pre_kernel<double><<<g, b, 0, st>>>(P, direct ? 1 : 0, (const double*)f, (const double*)f0,
                                                 (const double*)h, (const double*)h0, f1o, f0o, h1o, h0o);
    else
        read_pre_kernel<float><<<g, b, 0, st>>>(P, direct ? 1 : 0, (const float*)f, (const float*)f0,
                                                (const float*)h, (const float*)h0, f1o, f0o, h1o, h0o);
    return hipGetLastError();
}

hipError_t launch_write_pre(Prec prec, const LbmParams& P, const double* f1, const double* f0,
                            const double* h1, const double* h0, void* f, void* f0o, void* h, void* h0o,
                            hipStream_t st)
{
    dim3 g = grid_for(P.S), b(256);
    if (prec == Prec::F64)
        write_pre_kernel<double><<<g, b, 0, st>>>(P, f1, f0, h1, h0, (double*)f, (double*)f0o, (double*)h,
                                                  (double*)h0o);
    else
        write_pre_kernel<float><<<g, b, 0, st>>>(P, f1, f0, h1, h0, (float*)f, (float*)f0o, (float*)h,
                                                 (float*)h0o);
    return hipGetLastError();
}

}  // namespace ehd
