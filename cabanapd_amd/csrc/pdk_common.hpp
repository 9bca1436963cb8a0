// pdk_common.hpp -- device-side data layout shared by the engine's kernels.
//
// Layout in HBM (one specimen):
//   particles in lattice-row order (x fastest) when the positions form a
//   lattice, else in Morton order of their reference positions, SoA:
//     x[3][n] (FP64), U0[3][n] / U1[3][n] (FP64 displacement, double
//     buffered: the force pass reads U[cur] and its fused velocity-Verlet
//     epilogue writes U[cur^1]), v[3][n], f[3][n], vol[n], rho[n],
//     nofail[n] (u8), bcflag[n] (u8), optional body[3][n],
//     W[n], phi[n], m[n], theta[n].
//   bond family as a sliced CSR with 64 rows per slice (one wavefront per
//   slice, one lane per particle):
//     slice_ptr[nslices+1] (i64, in "slot rows": 64 bonds each),
//     cnt[n] (i32 row lengths),
//     col[total_slots][64] (i32 sorted-order neighbor index; padding = self),
//     intact[total_slots] (u64, bit lane = bond (row, slot) intact),
//     sid[total_slots][64] (u8 lattice stencil id, optional).
//   Row order within a slot row is ascending caller id, so a lane's serial
//   sum visits bonds in the oracle's order.
//   Stencil-slot form (full lattices): every slice has S slot rows, slot q =
//   stencil entry q for every lane, plus exist[total_slots] (u64, bit lane =
//   bond exists); the neighbor id is p + soff[q], so the force pass reads no
//   col stream.
#ifndef PDK_COMMON_HPP
#define PDK_COMMON_HPP

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pdk.h"

namespace pdk
{
constexpr int kSlice = 64;           // rows per slice = wavefront size
constexpr int kSlicesPerBlock = 4;   // 256-thread workgroups
constexpr int kBlock = kSlice * kSlicesPerBlock;
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, round-robin block dispatch

enum EpilogueMode : int32_t
{
    kEpiStoreOnly = 0,   // computeForce: store f (after BC/body terms)
    kEpiFinal = 1,       // final half step (v += dt/2 f/rho), store f
    kEpiFinalInitial = 2, // final half step of this step + initial half step of the next
    kEpiApplyOnly = 3     // store f after BC/body terms; the final half step runs as its own pass
};

// Layout of an entry of the FP64 stencil record (PDArgs::stencil_xiv): xi_x
// and |xi|^2 first, so the row-window pass loads the two values a slot needs
// with one scalar load (xi_y and xi_z are per lattice row).
constexpr int kXiX = 0, kXi2 = 1, kXiY = 2, kXiZ = 3;
// One 48-byte record per stencil entry for the row-window passes, so that a
// slot's values come in one scalar load: xi_x and |xi|^2 (FP64), then four
// FP32 values -- |xi|, infl |xi|, infl |xi|^2, 1/|xi| (the stencil table's w
// and the aux table's z, w, x) -- then xi_y and xi_z (FP64, read per row).
constexpr int kRecStride = 6, kRecXiX = 0, kRecXi2 = 1, kRecF = 2, kRecXiY = 4, kRecXiZ = 5;

// Everything a PD/contact kernel needs for one specimen.  Lives in device
// memory (one per spec, or an array of them for a packed batch).
struct PDArgs
{
    int64_t n;
    int64_t nslices;
    const double* x;        // [3][n]
    double* U[2];           // [3][n] each
    double* v;              // [3][n]
    double* f;              // [3][n]
    const double* vol;
    const double* rho;
    const uint8_t* nofail;
    const uint8_t* bcflag;
    const double* body;     // [3][n] or null
    const int64_t* slice_ptr;
    const int32_t* cnt;
    const int32_t* col;
    uint64_t* intact;
    const uint8_t* sid;     // stencil ids or null
    const float4* stencil;  // stencil table (xi_x, xi_y, xi_z, |xi|) or null
    // Stencil-slot form (full lattices): slot q of every slice is stencil
    // entry q; neighbor j = p + soff[q]; bit lane of exist[slice * S + q] says
    // whether the bond exists.  The stencil table is then indexed by q.
    // Null in the compact form, where slot q holds a bond iff q < cnt[p].
    const uint64_t* exist;
    const int32_t* soff;
    // Mixed form, per stencil entry in FP32: (1/|xi|, |xi|^2, w, w2) with the
    // LPS influence products w = infl |xi| (1 for influence 1, |xi| for
    // influence 0; the force coefficient's factor) and w2 = infl |xi|^2 (the
    // weighted-volume and dilatation factor)
    const float4* stencil_aux;
    const double* stencil_xi2d; // |xi|^2 of each stencil entry in FP64 (exact lattice offsets)
    const double* stencil_xiv;  // lattice stencil: 4 FP64 values per entry at kXiX, kXi2, kXiY, kXiZ
    const double* stencil_rec;  // stencil-slot form: kRecStride values per entry (kRec* layout)
    double* m;
    double* theta;
    double* lpsA;           // LPS: theta_coeff * theta / m per particle
    double* lpsB;           // LPS: s_coeff / m per particle
    double* W;
    double* phi;
    double* stress;         // [9][n], component 3 a + b (null until the first stress pass)
    const double* radius;   // contact: per-particle radius or null
    double* u_neigh;        // contact: u at last list build [3][n]
    double* block_partial;  // energy / max-displacement partials
    // model constants
    double c, bcc, s0, theta_coeff, s_coeff;
    double dt, half_dt;
    double bc_value;
    // mixed-precision guard band: a bond whose FP32 stretch lies within
    // [s0 - bandK, s0 + bandK] is re-tested in FP64 (bandf: the relative
    // width the gather passes scale by the bond's own terms)
    float s0f, bandf, bandK;
    // mixed form, uniform volume: the per-bond coefficient factor applied to
    // each FP32 row sum in FP64 (PMB: c vol, LPS: vol)
    double row_scale;
    // contact constants
    double Rs, Es, coeff_h_n, coeff_h_d, Gamma, a0, delta_tear, two_r, mu_f, nu, E;
    double jkr_c6; // pow(6, -1/3) (host libm): JKR tear-off of polydisperse pairs
    int32_t model;
    int32_t fracture;
    int32_t influence;
    int32_t precision;
    int32_t bc_kind;        // -1 none
    int32_t polydisperse;
    int32_t parity_offset;  // batch: current U index = launch parity ^ parity_offset
    int32_t unroll;         // force pass: bond slots loaded per group (1, 2, 4, 8)
    int64_t end_launch;     // batch: launch index of this spec's last step
    const int64_t* launch_ctr; // batch: the launch index on the device (graph replays pass launch < 0)
    // Exact shortcuts: every vol[j] equals vol_value / no particle is no-fail.
    int32_t vol_uniform;
    int32_t any_nofail;
    double vol_value;
    int32_t stencil_slots;  // S in the stencil-slot form, 0 in the compact form
    // Row windows of the stencil-slot form: the slots of one lattice row
    // (dy, dz) are consecutive and their neighbors of a slice form one
    // contiguous id window.  rows[r] = (first slot, slot count, id offset of
    // the first slot's neighbor, packed (dx0, dy, dz) + 128 per byte).
    const int4* rows;
    const uint32_t* rowstart; // bit q set: slot q starts a row (9 words), or null
    int32_t nrows;          // 0: no row table
    int32_t lat_n[3];       // lattice extent (stencil-slot form)
    // Multi-material specimens (FP64 gather passes): material of each
    // particle (null: one material), the PMB coefficient and the critical
    // stretch coefficient of each material pair [4 a + b], and each
    // material's LPS theta / s coefficients.
    const uint8_t* mtype;
    int32_t nmat;
    double mat_c[16], mat_bcc[16], mat_tc[4], mat_sc[4];
    // mixed form, several materials: the pair's critical stretch in FP32 and
    // its guard band (as s0f and bandK for one material)
    float mat_s0f[16], mat_bandK[16];
    // PMB family extensions (FP64 gather passes).  EPP: plastic stretch of
    // bond slot q of lane L of a slice at splas[(slice_ptr[slice] + q) *
    // kSlice + L] (null: Elastic) and the yield stretch sY.
    // TemperatureDependent: temperature per particle (null: independent),
    // expansion coefficient and reference temperature.
    double* splas;
    double sY;
    double* temp;
    double alpha, temp0;
    // DynamicTemperature: conduction per particle (heat pass scratch),
    // microconductivity constant, cp, horizon, linear microconductivity flag
    double* cond;
    double thermal_coeff, cp, thermal_horizon;
    int32_t lin_micro;
};

constexpr int kRowMaxSlots = 9;   // slots per row window (|dx| <= 4)
constexpr int kMaxRows = 64;      // rows per stencil
// Force-pass variant selected by the tuning knob: 0 = row windows staged in
// LDS (needs a row table), 1/2/4/8 = per-slot gathers in groups of that size.
__host__ __device__ __forceinline__ bool useRowWindows( const PDArgs& a )
{
    return a.unroll == 0 && a.rows != nullptr && a.nrows > 0;
}

// Bond-slot walk shared by both layouts: a lane visits slots [0, slotEnd) and
// slot q holds one of its bonds iff slotBond (q < cnt in the compact form,
// the exist bit in the stencil-slot form).  row0 = slice_ptr[slice].
__device__ __forceinline__ int32_t slotEnd( const PDArgs& A, int32_t len, int32_t mycnt )
{
    return A.exist ? len : mycnt;
}
__device__ __forceinline__ bool slotBond( const PDArgs& A, int64_t row0, int32_t q, int lane, int32_t mycnt )
{
    return A.exist ? ( ( ( A.exist[row0 + q] >> lane ) & 1ull ) != 0 ) : ( q < mycnt );
}

// Block -> logical slice group, keeping consecutive slice groups on one XCD
// so that their shared halo stays in that XCD's L2.
__device__ __forceinline__ int64_t xcdRemap( int64_t b, int64_t nb )
{
    const int64_t xcd = b % kNumXcd;
    const int64_t idx = b / kNumXcd;
    const int64_t q = nb / kNumXcd;
    const int64_t r = nb % kNumXcd;
    return ( xcd < r ) ? xcd * ( q + 1 ) + idx : r * ( q + 1 ) + ( xcd - r ) * q + idx;
}

__host__ __device__ __forceinline__ uint64_t spreadBits3( uint64_t v )
{
    v &= 0x1fffff;
    v = ( v | ( v << 32 ) ) & 0x1f00000000ffffULL;
    v = ( v | ( v << 16 ) ) & 0x1f0000ff0000ffULL;
    v = ( v | ( v << 8 ) ) & 0x100f00f00f00f00fULL;
    v = ( v | ( v << 4 ) ) & 0x10c30c30c30c30c3ULL;
    v = ( v | ( v << 2 ) ) & 0x1249249249249249ULL;
    return v;
}

__host__ __device__ __forceinline__ uint64_t morton3( uint64_t a, uint64_t b, uint64_t c )
{
    return spreadBits3( a ) | ( spreadBits3( b ) << 1 ) | ( spreadBits3( c ) << 2 );
}

} // namespace pdk

#endif
