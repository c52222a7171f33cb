// lbm_f32s.hip -- fused pull-scheme LBM step with FP32 storage of shifted
// populations (the performance path).
//
// Storage: g_d = f_d - w_d*rs(z) and k_d = h_d - w_d*qs(z), float, in row
// blocks [z][y][d-1][x] (the 26 moving populations of one x-row together;
// rest populations [z][y][x] apart), double buffered, where rs(z), qs(z) are the hydrostatic
// density and charge profiles of the box (fixed at creation).  The shifts take
// the large, smooth part of every population out of the float rounding: what
// is stored is the deviation from the hydrostatic state, so the rounding of
// a stored population is relative to that deviation, not to w*rho0 or to the
// hydrostatic pressure or charge.  A population pulled from plane z' is
// re-shifted to plane z by adding w_d (rs(z') - rs(z)).
//
// Same wall and representation rules as the FP64 kernel (lbm_f64.hip), which
// restates the reference step (LBM.cu:278-1353) in its own operation order.
// Here the collision is written directly on the deviations, pair by pair
// (d, d+1 are opposites):
//   rho = rs + sum g,  q = qs + sum k,
//   j_a = sum_d c_da g_d,  u = (j/CFL + F dt/2)/rho,  F = q E (+ exf x),
//   t_d = c_d.u /(CFL cs^2),
//   geq+ = w (rho - rs + rho (t^2/2 - u^2/(2 cs^2))),  geq- = w rho t,
//   F+  = w (c.u c.F /(CFL^2 cs^4) - u.F/cs^2),       F-  = w c.F /(CFL cs^2),
//   g'_d = g_d - om+ (g+ - geq+) - om- (g- - geq-) + dt (s+ F+ + s- F-),
// which is the algebra of gpu_collide_save (LBM.cu:519-1089) with the shift
// cancelling in every difference (w_d = w_opp(d), sum_d w_d = 1).  h uses
// u + K E, no source, and the charge rates.  Arithmetic is FP32.
#include <hip/hip_runtime.h>

#include <cstring>

#include "ehd_internal.h"

namespace ehd {
namespace {

// The float constants keep the lattice's exact identities.  The D3Q27
// weights 8/27, 2/27, 1/54, 1/216 share one mantissa, so in float all four
// are too large by the same 7.5e-9: sum w = 1 + 7.5e-9 and, with A rounded
// on its own, A sum_d w_d c_dz^2 = CFL (1 - 1.5e-8).  An equilibrium built on
// them conserves neither the density deviation nor the momentum: every
// collision scales the density deviation by 1 + om+ 7.5e-9 and the momentum
// by 1 - om- 1.5e-8, which on a plane-uniform wave (every node of a plane
// rounds alike, nothing averages out) grows into a relative velocity error of
// ~4e-6 per 1000 steps against the FP64 path.  So w0 = 1 - 6 ws - 12 wa - 8 wd
// of the float ws, wa, wd, and A = CFL / sum w c_z^2 of the same floats, are
// carried as hi + lo float pairs (one extra FMA where they are used): mass
// and momentum of the equilibrium then hold to double rounding.
struct F32Consts {
    float dt, cflinv, A, Alo; // A = 1/(CFL cs^2) = CFL / sum_d w_d c_dz^2 (hi + lo)
    float w0lo;               // w[0] + w0lo = 1 - 6 w[1] - 12 w[2] - 8 w[3]
    float inv2cs2;            // 1/(2 cs^2)
    float B;                  // 1/(CFL^2 cs^4)
    float ics2;               // 1/cs^2
    float K, exf;
    float omp, omm, ocp, ocm; // omega*dt
    float sp, sm;             // 1 - omega dt/2
    float w[4];
    float multi[4];
    float injk[4];            // 2 w (charge0 - qs(0))
    float hdx, hdy, hdz;      // 0.5/dx ...
};

// The buffer pointers of a box come from the item table in memory, so the
// compiler cannot see that they are global: without the address space it
// emits flat loads and stores.  Every access goes through these types.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const float cgfloat;
typedef __attribute__((address_space(1))) const double cgdouble;
__device__ __forceinline__ cgfloat* gc(const void* p) { return (cgfloat*)p; }
__device__ __forceinline__ gfloat* gw(void* p) { return (gfloat*)p; }
__device__ __forceinline__ cgdouble* gd(const double* p) { return (cgdouble*)p; }

// shift correction of a population of direction d pulled from plane z - c_z
__device__ __forceinline__ float pulled(float v, int d, float w, float dup, float ddn)
{
    if (CZ[d] > 0) return v + w * dup;
    if (CZ[d] < 0) return v + w * ddn;
    return v;
}

// Every population is read once and written once per step: streaming
// (non-temporal) loads and stores keep them from displacing the potential
// and the shift tables in the caches.
__device__ __forceinline__ float ldnt(cgfloat* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stnt(float v, gfloat* p) { __builtin_nontemporal_store(v, p); }

// One node s of one box: pull, collide, boundary and injection.
//
// TARGET (consistent start, after Mei et al., J. Comput. Phys. 2006): when
// the box carries target fields tg = [ux, uy, uz, q] (4 S doubles), every
// interior node collides towards the equilibria of the target velocity and
// charge instead of its own (its density stays free), the force uses the
// target charge, and the target charge is what the Poisson solve sees.
// Iterated, this builds the non-equilibrium parts of the populations that
// belong to the target state; the equilibrium start alone lacks them.
template <bool FIRST, bool EBUF, bool TARGET>
__device__ __forceinline__ void lbm_f32s_node(const LbmParams& P, const F32Consts& c, const PopBufs& pb,
                                              const MacroBufs& mb, const double* tgp, int macro, unsigned s)
{
    const unsigned nx = P.nx, ny = P.ny, nz = P.nz, nxy = P.nxy, S = P.S;
    const unsigned x = s % nx, t = s / nx, y = t % ny, z = t / ny;
    const unsigned xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
    const unsigned ym = (y == 0) ? ny - 1 : y - 1, yp = (y == ny - 1) ? 0 : y + 1;
    const unsigned zm = (z == 0) ? nz - 1 : z - 1, zp = (z == nz - 1) ? 0 : z + 1;
    const unsigned XS[3] = {xm, x, xp};
    // populations: row blocks [z][y][d-1][x] (26 x nx floats per row); the
    // source row of a pull along c is (y - c_y, z - c_z)
    const unsigned RW = 26u * nx;
    const unsigned YR[3] = {ym, y, yp}, ZR[3] = {zm * ny, z * ny, zp * ny};
    const unsigned own = t * RW + x;

    cgfloat* __restrict__ sh = gc(P.shf);
    const float rs = sh[z], dRu = sh[nz + z], dRd = sh[2 * nz + z];
    const float qs = sh[3 * nz + z], dQu = sh[4 * nz + z], dQd = sh[5 * nz + z];

    cgfloat* __restrict__ fA = gc(pb.f);
    cgfloat* __restrict__ f0A = gc(pb.f0);
    cgfloat* __restrict__ hA = gc(pb.h);
    cgfloat* __restrict__ h0A = gc(pb.h0);

    float g[27], k[27];
    g[0] = f0A[s];
    if (FIRST) {
        k[0] = h0A[s];
#pragma unroll
        for (int d = 1; d < 27; ++d) {
            g[d] = fA[own + (d - 1) * nx];
            k[d] = hA[own + (d - 1) * nx];
        }
    } else {
#pragma unroll
        for (int d = 1; d < 27; ++d)
            g[d] = pulled(ldnt(fA + (ZR[1 - CZ[d]] + YR[1 - CY[d]]) * RW + (d - 1) * nx + XS[1 - CX[d]]), d,
                          c.w[WCLS[d]], dRu, dRd);
        if (z == 0) {
            k[0] = -h0A[s] + c.injk[0];
#pragma unroll
            for (int d = 1; d < 27; ++d) k[d] = -hA[own + (opp(d) - 1) * nx] + c.injk[WCLS[d]];
        } else {
            k[0] = h0A[s];
#pragma unroll
            for (int d = 1; d < 27; ++d)
                k[d] = pulled(ldnt(hA + (ZR[1 - CZ[d]] + YR[1 - CY[d]]) * RW + (d - 1) * nx + XS[1 - CX[d]]), d,
                              c.w[WCLS[d]], dQu, dQd);
        }
    }

    // E = -grad(phi): float fluctuation plus FP64 plane means
    cgfloat* __restrict__ phi = gc(mb.phi);
    cgdouble* __restrict__ phibar = gd(mb.phibar);
    float Ex, Ey, Ez;
    const unsigned yc = y * nx, zc = z * nxy;
    if (EBUF) {
        Ex = (float)gd(mb.ex)[s]; Ey = (float)gd(mb.ey)[s]; Ez = (float)gd(mb.ez)[s];
    } else {
        Ex = c.hdx * (phi[xm + yc + zc] - phi[xp + yc + zc]);
        Ey = c.hdy * (phi[x + ym * nx + zc] - phi[x + yp * nx + zc]);
        const unsigned ze = (z == 0) ? 1u : ((z == nz - 1) ? nz - 2 : z);
        const float dmn = (float)(phibar[ze - 1] - phibar[ze + 1]);
        Ez = c.hdz * ((phi[x + yc + (ze - 1) * nxy] - phi[x + yc + (ze + 1) * nxy]) + dmn);
    }

    // moments
    float drho = g[0], dq = k[0];
#pragma unroll
    for (int d = 1; d < 27; d += 2) {
        drho += g[d] + g[d + 1];
        dq += k[d] + k[d + 1];
    }
    const float rho = rs + drho;
    const float charge = qs + dq;
    // charge of the equilibrium and of the force (the node's own unless TARGET)
    float qe = charge, dqe = dq;
    bool tnode = false;
    if constexpr (TARGET) {
        if (tgp && z > 0 && z < nz - 1) {
            tnode = true;
            qe = (float)gd(tgp)[3 * (size_t)S + s];
            dqe = qe - qs;
        }
    }
    float j[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float pp = 0.f, nn = 0.f;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            pp += g[MOM[a][i]];
            nn += g[MOM[a][9 + i]];
        }
        j[a] = pp - nn;
    }
    const float Fx = qe * Ex + c.exf, Fy = qe * Ey, Fz = qe * Ez;
    const float hdt = 0.5f * c.dt;
    const float rinv = 1.0f / rho;
    float ux = rinv * (j[0] * c.cflinv + Fx * hdt);
    float uy = rinv * (j[1] * c.cflinv + Fy * hdt);
    float uz = rinv * (j[2] * c.cflinv + Fz * hdt);

    if (z == 0) {
        // velocity override from the pre-collision populations of z=1 (LBM.cu:416-493)
        float gm[27];
        const unsigned s1 = s + nxy;
        float dqm = h0A[s1];
        if (FIRST) {
            const unsigned own1 = own + ny * RW;
#pragma unroll
            for (int d = 1; d < 27; ++d) {
                gm[d] = fA[own1 + (d - 1) * nx];
                dqm += hA[own1 + (d - 1) * nx];
            }
        } else {
            const float dRu1 = sh[nz + 1], dRd1 = sh[2 * nz + 1], dQu1 = sh[4 * nz + 1], dQd1 = sh[5 * nz + 1];
#pragma unroll
            for (int d = 1; d < 27; ++d) {
                // node (x, y, 1) pulls from row (y - c_y, 1 - c_z)
                const unsigned src = ((unsigned)(1 - CZ[d]) * ny + YR[1 - CY[d]]) * RW + (d - 1) * nx + XS[1 - CX[d]];
                gm[d] = pulled(fA[src], d, c.w[WCLS[d]], dRu1, dRd1);
                dqm += pulled(hA[src], d, c.w[WCLS[d]], dQu1, dQd1);
            }
        }
        float Exm, Eym, Ezm;
        if (EBUF) {
            Exm = (float)gd(mb.ex)[s1]; Eym = (float)gd(mb.ey)[s1]; Ezm = (float)gd(mb.ez)[s1];
        } else {
            Exm = c.hdx * (phi[xm + yc + nxy] - phi[xp + yc + nxy]);
            Eym = c.hdy * (phi[x + ym * nx + nxy] - phi[x + yp * nx + nxy]);
            const float dmn = (float)(phibar[0] - phibar[2]);
            Ezm = c.hdz * ((phi[x + yc] - phi[x + yc + 2 * nxy]) + dmn);
        }
        const float qm = sh[3 * nz + 1] + dqm;
        float jm[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float pp = 0.f, nn = 0.f;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                pp += gm[MOM[a][i]];
                nn += gm[MOM[a][9 + i]];
            }
            jm[a] = pp - nn;
        }
        ux = -rinv * (jm[0] * c.cflinv + (qm * Exm + c.exf) * hdt);
        uy = -rinv * (jm[1] * c.cflinv + (qm * Eym) * hdt);
        uz = -rinv * (jm[2] * c.cflinv + (qm * Ezm) * hdt);
    }

    if constexpr (TARGET) {
        if (tnode) {
            cgdouble* tg = gd(tgp);
            ux = (float)tg[s];
            uy = (float)tg[(size_t)S + s];
            uz = (float)tg[2 * (size_t)S + s];
        }
    }

    if (macro) {
        gw(mb.rho)[s] = rho;
        gw(mb.ux)[s] = ux;
        gw(mb.uy)[s] = uy;
        gw(mb.uz)[s] = uz;
    }
    gw(mb.q)[s] = qe;

    const float ucx = ux + c.K * Ex, ucy = uy + c.K * Ey, ucz = uz + c.K * Ez;
    const float U2 = (ux * ux + uy * uy + uz * uz) * c.inv2cs2;
    const float Uc2 = (ucx * ucx + ucy * ucy + ucz * ucz) * c.inv2cs2;
    const float uF = (ux * Fx + uy * Fy + uz * Fz) * c.ics2;

    gfloat* __restrict__ fB = gw(pb.fo);
    gfloat* __restrict__ f0B = gw(pb.f0o);
    gfloat* __restrict__ hB = gw(pb.ho);
    gfloat* __restrict__ h0B = gw(pb.h0o);
    const bool fwall = (z == 0) || (z == nz - 1);

    // rest populations
    {
        const float xg = drho - rho * U2, xk = dqe - qe * Uc2;
        const float geq0 = fmaf(c.w[0], xg, c.w0lo * xg);
        const float keq0 = fmaf(c.w[0], xk, c.w0lo * xk);
        const float F0 = -fmaf(c.w[0], uF, c.w0lo * uF);
        const float g0n = g[0] - c.omp * (g[0] - geq0) + c.dt * c.sp * F0;
        const float k0n = k[0] - c.ocp * (k[0] - keq0);
        f0B[s] = fwall ? g[0] : g0n;
        if (z != nz - 1) {
            h0B[s] = k0n;
            if (z == nz - 2) h0B[s + nxy] = k0n;   // qs(nz-1) = qs(nz-2): no re-shift
        }
    }
#pragma unroll
    for (int d = 1; d < 27; d += 2) {
        const int e = d + 1;
        const float w = c.w[WCLS[d]];
        const float cu = (float)CX[d] * ux + (float)CY[d] * uy + (float)CZ[d] * uz;
        const float cc = (float)CX[d] * ucx + (float)CY[d] * ucy + (float)CZ[d] * ucz;
        const float cF = (float)CX[d] * Fx + (float)CY[d] * Fy + (float)CZ[d] * Fz;
        const float tt = fmaf(c.A, cu, c.Alo * cu), tc = fmaf(c.A, cc, c.Alo * cc);
        const float geqp = w * (drho + rho * (0.5f * tt * tt - U2));
        const float geqm = w * rho * tt;
        const float keqp = w * (dqe + qe * (0.5f * tc * tc - Uc2));
        const float keqm = w * qe * tc;
        const float Fp = w * (c.B * cu * cF - uF);
        const float Fm = w * fmaf(c.A, cF, c.Alo * cF);
        const float gp = 0.5f * (g[d] + g[e]), gmn = 0.5f * (g[d] - g[e]);
        const float kp = 0.5f * (k[d] + k[e]), kmn = 0.5f * (k[d] - k[e]);
        const float rp = c.omp * (gp - geqp), rm = c.omm * (gmn - geqm);
        const float sP = c.dt * c.sp * Fp, sM = c.dt * c.sm * Fm;
        const float qp = c.ocp * (kp - keqp), qm2 = c.ocm * (kmn - keqm);
        const float gd = g[d] - rp - rm + sP + sM;
        const float ge = g[e] - rp + rm + sP - sM;
        const float kd = k[d] - qp - qm2;
        const float ke = k[e] - qp + qm2;
        // f: bounce-back of the pre-collision values at the walls (same plane, same shift)
        float od, oe;
        if (z == 0) {
            od = g[e];
            oe = g[d];
        } else if (z == nz - 1) {
            od = g[e] + (float)TOPSGN[d] * c.multi[WCLS[d]];
            oe = g[d] + (float)TOPSGN[e] * c.multi[WCLS[d]];
        } else {
            od = gd;
            oe = ge;
        }
        stnt(od, fB + own + (d - 1) * nx);
        stnt(oe, fB + own + (e - 1) * nx);
        if (z != nz - 1) {
            stnt(kd, hB + own + (d - 1) * nx);
            stnt(ke, hB + own + (e - 1) * nx);
            if (z == nz - 2) {
                stnt(kd, hB + own + ny * RW + (d - 1) * nx);
                stnt(ke, hB + own + ny * RW + (e - 1) * nx);
            }
        }
    }
}

// One launch over one box or a batch of boxes: the boxes' node ranges are
// cut into 256-node tiles numbered across the batch (box i owns tiles
// tile0[i] .. tile0[i+1]-1) and the workgroups stride over all tiles (one
// tile each when the grid covers them, persistent otherwise).  A single box
// is a batch of one, so a box gives the same numbers alone and in a batch.
//
// Tile order: the hardware deals workgroup b to XCD b mod 8, and each XCD
// has its own L2.  With chunk > 0 the tiles are taken in superblocks of
// 8*chunk (one z-plane of a box when chunk = tiles per plane / 8): XCD j
// gets the j-th run of chunk consecutive tiles of every superblock, so the
// potential rows a tile reads around itself (y +- 1, z +- 1) are mostly the
// rows its own XCD reads as well and come from that L2 instead of HBM.
// Every superblock still spreads over all XCDs, so the wall planes (more
// loads per node) do not pile up on one XCD.  chunk = 0 is plain linear
// order.  The grid is a multiple of 8 whenever it is persistent.
struct BatchItem {
    LbmParams P;
    F32Consts c;
    PopBufs pb;
    MacroBufs mb;
    const double* target;   // consistent-start target fields, or null
    unsigned tile0, ntiles;
};

template <bool FIRST, bool EBUF, bool TARGET>
__global__ void __launch_bounds__(256) lbm_f32s_kernel(const BatchItem* __restrict__ items, int nitems,
                                                             unsigned total_tiles, unsigned chunk, int macro)
{
    const unsigned sb = 8u * chunk;
    const unsigned full = chunk ? (total_tiles / sb) * sb : 0u;
    for (unsigned v = blockIdx.x; v < total_tiles; v += gridDim.x) {
        unsigned tile = v;
        if (v < full) {
            const unsigned k = v >> 3, j = v & 7u;
            tile = (k / chunk) * sb + j * chunk + k % chunk;
        }
        int lo = 0, hi = nitems - 1;   // last item with tile0 <= tile
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].tile0 <= tile) lo = mid;
            else hi = mid - 1;
        }
        const BatchItem& it = items[lo];
        const unsigned s = (tile - it.tile0) * 256u + threadIdx.x;
        if (s < it.P.S) lbm_f32s_node<FIRST, EBUF, TARGET>(it.P, it.c, it.pb, it.mb, TARGET ? it.target : nullptr,
                                                           macro, s);
    }
}

F32Consts make_f32consts(const LbmParams& P)
{
    F32Consts c;
    const double cs2 = P.cs_square, cfl = P.cfl;
    c.dt = (float)P.dt;
    c.cflinv = (float)(1.0 / cfl);
    {
        // identities of the float weights (see F32Consts)
        const double ws = (float)P.w[1], wa = (float)P.w[2], wd = (float)P.w[3];
        const double w0 = 1.0 - 6.0 * ws - 12.0 * wa - 8.0 * wd;
        const double A = cfl / (2.0 * ws + 8.0 * wa + 8.0 * wd);
        c.A = (float)A;
        c.Alo = (float)(A - (double)c.A);
        c.w0lo = (float)(w0 - (double)(float)w0);
    }
    c.inv2cs2 = (float)(0.5 / cs2);
    c.B = (float)(1.0 / (cfl * cfl * cs2 * cs2));
    c.ics2 = (float)(1.0 / cs2);
    c.K = (float)P.K;
    c.exf = (float)P.exf;
    c.omp = (float)(P.omega_plus * P.dt);
    c.omm = (float)(P.omega_minus * P.dt);
    c.ocp = (float)(P.omega_c_plus * P.dt);
    c.ocm = (float)(P.omega_c_minus * P.dt);
    c.sp = (float)(1.0 - 0.5 * P.dt * P.omega_plus);
    c.sm = (float)(1.0 - 0.5 * P.dt * P.omega_minus);
    for (int i = 0; i < 4; ++i) {
        c.w[i] = (i == 0) ? (float)(1.0 - 6.0 * (double)(float)P.w[1] - 12.0 * (double)(float)P.w[2] -
                                    8.0 * (double)(float)P.w[3])
                          : (float)P.w[i];
        c.multi[i] = (float)P.multi[i];
        c.injk[i] = (float)(2.0 * P.w[i] * (P.charge0 - P.qs0));
    }
    c.hdx = (float)(0.5 / P.dx);
    c.hdy = (float)(0.5 / P.dy);
    c.hdz = (float)(0.5 / P.dz);
    return c;
}

}  // namespace

size_t lbm_f32s_batch_item_size() { return sizeof(BatchItem); }

void lbm_f32s_batch_item(void* out, const LbmParams& P, const PopBufs& pb, const MacroBufs& mb, unsigned tile0,
                         const double* target)
{
    BatchItem it{};
    it.P = P;
    it.c = make_f32consts(P);
    it.pb = pb;
    it.mb = mb;
    it.target = target;
    it.tile0 = tile0;
    it.ntiles = (P.S + 255) / 256;
    std::memcpy(out, &it, sizeof(it));
}

template <bool TARGET>
static void launch_variant(const BatchItem* it, int nitems, unsigned total_tiles, unsigned g, unsigned chunk, bool first,
                    bool ebuf, int m, hipStream_t st)
{
    if (first && ebuf) lbm_f32s_kernel<true, true, TARGET><<<g, 256, 0, st>>>(it, nitems, total_tiles, chunk, m);
    else if (first) lbm_f32s_kernel<true, false, TARGET><<<g, 256, 0, st>>>(it, nitems, total_tiles, chunk, m);
    else if (ebuf) lbm_f32s_kernel<false, true, TARGET><<<g, 256, 0, st>>>(it, nitems, total_tiles, chunk, m);
    else lbm_f32s_kernel<false, false, TARGET><<<g, 256, 0, st>>>(it, nitems, total_tiles, chunk, m);
}

hipError_t launch_lbm_f32s_batch(const void* items, int nitems, unsigned total_tiles, unsigned grid, unsigned chunk,
                                 bool first, bool ebuf, bool macro, hipStream_t st, bool target)
{
    if (nitems < 1 || total_tiles == 0) return hipErrorInvalidValue;
    const BatchItem* it = (const BatchItem*)items;
    const unsigned g = (grid == 0 || grid > total_tiles) ? total_tiles : grid;
    if (g < total_tiles && g % 8u) return hipErrorInvalidValue;   // the tile order assumes b mod 8 = XCD
    const int m = macro ? 1 : 0;
    if (target) launch_variant<true>(it, nitems, total_tiles, g, chunk, first, ebuf, m, st);
    else launch_variant<false>(it, nitems, total_tiles, g, chunk, first, ebuf, m, st);
    return hipGetLastError();
}

unsigned lbm_xcd_chunk(unsigned nxy)
{
    if (nxy % 256u) return 0u;
    const unsigned tpp = nxy / 256u;   // tiles per plane
    return (tpp % 8u) ? 0u : tpp / 8u;
}


}  // namespace ehd
