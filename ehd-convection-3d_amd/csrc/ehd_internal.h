// ehd_internal.h -- types shared by the HIP kernels and the C-ABI host layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/ehd.h"
#include "ehd_lattice.h"

namespace ehd {

// Per-box constants handed to the kernels by value.
struct LbmParams {
    int nx, ny, nz;
    unsigned nxy, S;             // nx*ny, nx*ny*nz
    double dt, cfl, cs_square, K, exf, charge0, rho0;
    double w[4];                 // w0, ws, wa, wd (LBM.h:73-76)
    double omega_plus, omega_minus, omega_c_plus, omega_c_minus;  // LBM.cu:298-302
    double multi[4];             // moving-wall terms per weight class (LBM.cu:1137-1139)
    double inj[4];               // 2*charge0*w (LBM.cu:1288-1291)
    double dx, dy, dz;
    // FP32S storage shifts (nullptr for F64): f is stored as f - w*rs[z], h as
    // h - w*qs[z]; shf packs the float tables rs, rs[zm]-rs, rs[zp]-rs, qs,
    // qs[zm]-qs, qs[zp]-qs, nz entries each (zm, zp: periodic neighbours).
    const double* rs;
    const double* qs;
    const float* shf;
    double qs0;                  // host copy of qs[0]
};

// Device buffers of the Poisson solve.
struct PoissonBufs {
    int nx, ny, nz, nxh;         // nxh = nx/2+1
    void* spec;                  // complex [nz-2][ny][nxh], element type of the path
    const void* twx;             // nx roots exp(-2 pi i j/nx), element type of the path
    const void* twy;             // ny roots
    double* phibar;              // FP32S: xy-mean potential per plane [nz] (FP64)
};

enum class Prec { F64 = 0, F32S = 1 };

// --- launch wrappers (lbm_f64.hip) ---------------------------------------
struct PopBufs {
    void* f;  void* f0;  void* h;  void* h0;     // state read (A)
    void* fo; void* f0o; void* ho; void* h0o;    // state written (B)
};
struct MacroBufs {
    void* rho; void* ux; void* uy; void* uz; void* q;   // element type of the path
    const void* phi;                                     // potential (path type)
    const double* phibar;                                // FP32S plane means (nullptr for F64)
    const double* ex; const double* ey; const double* ez;   // explicit E (first step)
};

hipError_t launch_lbm_f64(const LbmParams& P, const PopBufs& pb, const MacroBufs& mb,
                          bool first, bool ebuf, bool macro, hipStream_t st);
// FP32S LBM over one box or a batch of boxes: an opaque item per box (lbm_f32s_batch_item writes
// lbm_f32s_batch_item_size() bytes), items in a device array sorted by tile0,
// tile0 = number of 256-node tiles of the boxes before it.  grid = 0 launches
// one workgroup per tile; a smaller grid (a multiple of 8) makes the
// workgroups persistent.  chunk: XCD tile order (0 = linear), see
// lbm_xcd_chunk.
size_t lbm_f32s_batch_item_size();
// target: consistent-start target fields [ux, uy, uz, q] (4 S doubles) of the
// box, used only by launches with target = true (null: the box steps normally).
void lbm_f32s_batch_item(void* out, const LbmParams& P, const PopBufs& pb, const MacroBufs& mb, unsigned tile0,
                         const double* target = nullptr);
hipError_t launch_lbm_f32s_batch(const void* items, int nitems, unsigned total_tiles, unsigned grid, unsigned chunk,
                                 bool first, bool ebuf, bool macro, hipStream_t st, bool target = false);
// XCD tile-order chunk for boxes of nxy nodes per plane: one plane per
// superblock when a plane is a whole number of 8-tile groups, else 0.
unsigned lbm_xcd_chunk(unsigned nxy);

// init_equilibrium (LBM.cu:98-276) from double macroscopic fields into the
// pre-collision ("direct") representation of the path.
hipError_t launch_init_equilibrium(Prec prec, const LbmParams& P, const double* rho, const double* q,
                                   const double* ux, const double* uy, const double* uz,
                                   const double* ex, const double* ey, const double* ez,
                                   void* f, void* f0, void* h, void* h0, hipStream_t st);

// Pre-collision populations of the reference (after stream + bc_charge) as
// doubles, from either representation.
hipError_t launch_read_pre(Prec prec, const LbmParams& P, bool direct, const void* f, const void* f0,
                           const void* h, const void* h0, double* f1o, double* f0o, double* h1o,
                           double* h0o, hipStream_t st);
hipError_t launch_write_pre(Prec prec, const LbmParams& P, const double* f1, const double* f0,
                            const double* h1, const double* h0, void* f, void* f0o, void* h, void* h0o,
                            hipStream_t st);

// E = -grad(phi) with the wall copy of Ez (poisson.cu:32-73), as doubles.
hipError_t launch_efield(Prec prec, const LbmParams& P, const void* phi, const double* phibar,
                         double* ex, double* ey, double* ez, hipStream_t st);

// Macroscopic field conversion helpers.
hipError_t launch_to_double(Prec prec, const void* src, double* dst, size_t n, hipStream_t st);
hipError_t launch_from_double(Prec prec, const double* src, void* dst, size_t n, hipStream_t st);
hipError_t launch_phi_full(Prec prec, const LbmParams& P, const void* phi, const double* phibar,
                           double* out, hipStream_t st);
hipError_t launch_phi_split(Prec prec, const LbmParams& P, const double* phi, void* phio,
                            double* phibar, hipStream_t st);

// Reference and hydrostatic initial fields (doubles).
hipError_t launch_init_fields(const LbmParams& P, int mode, const double* qz, const double* phiz,
                              const double* rhoz, uint64_t seed, double amp, int pattern, double lx, double ly,
                              double voltage, double lz, double* rho, double* q, double* phi, double* ux,
                              double* uy, double* uz, hipStream_t st);
int diag_blocks();

// --- Poisson (poisson.hip) -------------------------------------------------
// One box's share of a Poisson solve.  A batch runs many jobs of one (nx, ny)
// in the same launches (job index = blockIdx.y); a single box is a group of one.
struct PoissonJob {
    const void* q;               // charge (path type), [nz][ny][nx]
    void* spec;                  // spectrum workspace of the box
    void* phi;                   // potential (path type)
    double* phibar;              // FP32S plane means, nullptr for F64
    int n;                       // interior planes nz-2
    double kxs, kys, dz2, eps, voltage;
};
struct PoissonGroup {
    int nx, ny, max_n, njobs;
    int split_mean;              // FP32S
    int lpb, ylines;             // lines per block of the x / y passes (0: largest that fits)
    int zs_partition;            // threads per mode of the z solve (poisson_shape)
    const void* twx;             // twiddles of nx and ny, path type
    const void* twy;
    const PoissonJob* jobs;      // device array of njobs
    int debug_sync;              // synchronise the stream and check after every pass
};
// name of the Poisson pass whose synchronised check failed last (debug_sync)
const char* poisson_failed_pass();
struct PoissonShape {
    int max_x_lines, max_y_lines, zs_partition;
};
PoissonJob make_poisson_job(const LbmParams& P, const PoissonBufs& pb, const void* q, void* phi, double voltage,
                            double eps);
PoissonShape poisson_shape(Prec prec, int nx, int ny, int nz);
hipError_t launch_poisson_group(Prec prec, const PoissonGroup& g, hipStream_t st);

// --- diagnostics (diag.hip) --------------------------------------------------
// out[0]=sum_xy q_top*Ez_top, out[1]=sum interior q, out[2]=max uz (top extrapolated),
// out[3]=sum interior rho, out[4]=count of nodes with a non-finite q, rho, ux, uy or uz
// scratch: 5 * diag_blocks() doubles of partial results
hipError_t launch_diag(Prec prec, const LbmParams& P, const void* q, const void* ux, const void* uy, const void* uz,
                       const void* rho, const void* phi, const double* phibar, double* scratch, double* out,
                       hipStream_t st);

}  // namespace ehd
